"""GPU parity: libdgrain on gfx950, called through the C ABI, against the
oracle in its device-math mode (squares as x*x, H^(4/3) by the shared
deterministic algorithm).  The bar is bit-for-bit equality up to the sign of
zero, for states, RHS, edge fluxes, max_lambda, dt, wet/dry flags and limiter
activations.  (The oracle's reference-math mode is pinned to the compiled
reference in test_oracle_vs_reference.py; the two modes are compared to the
tolerance rel-L2 <= 1e-11 in test_oracle_modes.py.)"""
import os

import numpy as np
import pytest

import oracle
from dgrain import configs, native
from helpers import beach_mesh, beach_state, perturbed_mesh, random_state, same, rel_l2

pytestmark = pytest.mark.gpu


def _pair(mesh, wetdry=0, **kw):
    s = native.Solver(mesh, native.default_params(wetdry=wetdry, **kw))
    o = oracle.Oracle(mesh, math=oracle.MATH_DEV, wetdry=wetdry, **kw)
    return s, o


def test_math_selftest_matches_host(gpu):
    rng = np.random.default_rng(7)
    n = 200000
    x = np.concatenate([np.exp(rng.uniform(np.log(1e-6), np.log(1e4), n)),
                        rng.uniform(-50, 50, 1000), [0.0, -0.0, 1.0, 8.0, 1e-310, 5e-324, 1e300],
                        [np.inf, -np.inf, np.nan, -1.0, 2.2250738585072014e-308, 2.2250738585072009e-308,
                         1.7976931348623157e308]])
    y = rng.uniform(-3, 3, x.size)
    out = native.math_selftest(x, y)
    assert same(out[:, 0], np.sqrt(x))                  # IEEE sqrt
    assert same(out[:, 1], x / y)                        # IEEE division
    o43 = np.array([oracle.pow43(v) for v in x])
    assert same(out[:, 2], o43)
    assert np.isnan(out[-7:-4, 2]).sum() == 2 and out[-7, 2] == np.inf   # inf, -inf -> NaN, NaN
    osin = np.array([oracle.dsin(v) for v in x])
    assert same(out[:, 3], osin)


def _bits_equal(a, b):
    """Bitwise equality, the sign of zero included (any NaN == any NaN)."""
    return bool(np.all((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))))


def test_shared_denominator_division(gpu):
    """quot(x, divisor(y)) is x / y bit for bit, on ordinary operands (the
    3-instruction path) and on zeros, denormals, huge values, inf and NaN
    (the fallback), including values around the fallback's thresholds."""
    rng = np.random.default_rng(11)
    n = 2_000_000
    sx = rng.choice([-1.0, 1.0], n)
    sy = rng.choice([-1.0, 1.0], n)
    x = sx * np.exp2(rng.uniform(-60, 60, n)) * rng.uniform(1, 2, n)
    y = sy * np.exp2(rng.uniform(-60, 60, n)) * rng.uniform(1, 2, n)
    # operands whose quotient or divisor straddles 2^-300, 2^300, 2^-320, 2^320
    m = 200_000
    ex = rng.choice([-330.0, -320.0, -310.0, -300.0, -290.0, 290.0, 300.0, 310.0, 320.0, 330.0], m)
    x2 = np.exp2(ex + rng.uniform(-2, 2, m)) * rng.uniform(1, 2, m)
    y2 = np.exp2(rng.choice([-330.0, -320.0, -10.0, 0.0, 10.0, 320.0, 330.0], m) + rng.uniform(-2, 2, m))
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 1e-310, 1e308, -1e308,
                        1.0, 3.0, 1.5, 2.0 ** 1023, 2.0 ** -1022])
    xs, ys = np.meshgrid(special, special)
    x = np.concatenate([x, x2, y2, xs.ravel(), rng.uniform(-1e4, 1e4, 1000)])
    y = np.concatenate([y, y2, x2, ys.ravel(), rng.uniform(-1e4, 1e4, 1000)])
    out = native.math_selftest(x, y)
    with np.errstate(all="ignore"):
        q = x / y
        q3 = x / 3.0
    assert _bits_equal(out[:, 1], q)
    assert _bits_equal(out[:, 5], q)
    assert _bits_equal(out[:, 6], q3)


def test_green_ampt_math_matches_oracle(gpu):
    """The device's ln(1 + x) and Green-Ampt depth F(t) against the oracle's,
    bit for bit, over the time range a run meets (and t = 0, denormals)."""
    rng = np.random.default_rng(12)
    x = np.concatenate([np.exp(rng.uniform(np.log(1e-9), np.log(1e8), 20000)), [0.0, 5e-324, 1e-310, 1.0, 2.0]])
    y = rng.uniform(-3, 3, x.size)
    out = native.math_selftest(x, y)
    l1p = np.array([oracle.dlog1p(abs(v)) for v in x])
    assert _bits_equal(out[:, 7], l1p)
    F = np.array([oracle.ga_cumulative(1e-6 + 1e-5 * abs(b), 0.033, abs(a)) for a, b in zip(x, y)])
    assert _bits_equal(out[:, 8], F)


@pytest.mark.parametrize("wetdry", [0, 1])
@pytest.mark.parametrize("seed", [1, 2])
def test_compute2DL(gpu, wetdry, seed):
    mesh = perturbed_mesh(N=8, seed=seed)
    s, o = _pair(mesh, wetdry)
    z, qx, qy = random_state(mesh, seed=seed + 10)
    s.set_state(z, qx, qy)
    o.set_state(z, qx, qy)
    if wetdry:
        wd = np.random.default_rng(seed).integers(0, 2, mesh.num_el).astype(np.int32)
        s.set_wd(wd)
        o.set_wd(wd)
    rg, fg, lg = s.compute2DL()
    ro, fo, lo = o.compute2DL()
    for a, b in zip(fg, fo):
        assert same(a, b)
    for a, b in zip(rg, ro):
        assert same(a, b)
    assert lg == lo


def _dry_state(mesh, dry, seed):
    """Random state whose `dry` elements hold still water: zero depth (zeta =
    -z_edge) at the midpoints of edges shared with another dry element or on
    the boundary, 5 cm elsewhere.  The both-dry edges then give NaN
    max-|lambda| operands (0/0 velocities), which restart the reference's
    fold, while every wet/dry edge stays finite."""
    z, qx, qy = random_state(mesh, seed=seed)
    ev = np.asarray(mesh.elto_vert).reshape(-1, 3)
    zn = np.asarray(mesh.z)
    owner = {}
    for e in range(ev.shape[0]):
        for k in range(3):
            owner.setdefault(frozenset((ev[e, k], ev[e, (k + 1) % 3])), []).append(e)
    for e in np.flatnonzero(dry):
        for k in range(3):
            a, b = ev[e, k], ev[e, (k + 1) % 3]
            els = owner[frozenset((a, b))]
            other = [x for x in els if x != e]
            ze = (zn[a] + zn[b]) / 2
            z[3 * e + k] = -ze if (not other or dry[other[0]]) else 0.05 - ze
            qx[3 * e + k] = 0.0
            qy[3 * e + k] = 0.0
    return z, qx, qy


@pytest.mark.parametrize("variant", range(native.NUM_VARIANTS))
@pytest.mark.parametrize("pattern", ["band", "tail", "all"])
def test_max_lambda_fold_with_nan_operands(gpu, variant, pattern):
    """max |lambda| when NaN operands occur (both-dry edges at zero depth):
    the fold restarts after each NaN, so the result is the max of the operands
    after the last one (band: NaNs in the middle of the edge order; tail: NaNs
    up to the last few edges, whose operands alone give the result; all: every
    operand NaN, result NaN), for every kernel variant, against the oracle's
    left fold."""
    mesh = perturbed_mesh(N=12, seed=4)
    ne = mesh.num_el
    dry = np.zeros(ne, bool)
    if pattern == "band":
        dry[ne // 3: ne // 2] = True
    elif pattern == "tail":
        dry[ne - ne // 4: ne - 4] = True
    else:
        dry[:] = True
    s, o = _pair(mesh, 1)
    s.set_variant(variant)
    z, qx, qy = _dry_state(mesh, dry, seed=variant + 30)
    wd = np.where(dry, 0, 1).astype(np.int32)
    for obj in (s, o):
        obj.set_state(z, qx, qy)
        obj.set_wd(wd)
    rg, fg, lg = s.compute2DL()
    ro, fo, lo = o.compute2DL()
    assert same(np.array([lg]), np.array([lo])), (lg, lo)
    if pattern == "all":
        assert np.isnan(lg)
    else:
        assert np.isfinite(lg) and lg > 0
    for a, b in zip(fg, fo):
        assert same(a, b)


def test_limiter_wd_pdop(gpu):
    mesh = beach_mesh(N=12)
    s, o = _pair(mesh, 1)
    z, qx, qy = random_state(mesh, seed=5, dz=0.6, q=0.05)
    z = z + 0.2
    for obj in (s, o):
        obj.set_state(z, qx, qy)
        obj.wetDryStatus2D()
    assert np.array_equal(s.get_wd(), o.get_wd())
    s.SlopeLimiter()
    mo = o.SlopeLimiter()
    assert np.array_equal(s.get_modify(0), mo)
    for a, b in zip(s.get_state(), o.get_state()):
        assert same(a, b)
    # PDop on a state with negative vertex depths
    z2 = z - 0.3
    for obj in (s, o):
        obj.set_state(z2, qx, qy)
        obj.PDop2D()
    for a, b in zip(s.get_state(), o.get_state()):
        assert same(a, b)


def _compare_run(s, o, nsteps, chunk):
    done = 0
    while done < nsteps:
        k = min(chunk, nsteps - done)
        rep = s.run(k)
        t, dt = o.run(k)
        done += k
        for a, b in zip(s.get_state(), o.get_state()):
            assert same(a, b), "state differs after %d steps" % done
        assert np.array_equal(s.get_wd(), o.get_wd()), "WD differs after %d steps" % done
        for st in (0, 1):
            assert np.array_equal(s.get_modify(st), o.get_modify(st)), "limiter flags differ after %d steps" % done
        assert rep.t == t and rep.dt == dt, (done, rep.t, t, rep.dt, dt)
        print("  %d/%d steps bitwise equal (t=%.6g)" % (done, nsteps, t), flush=True)


def test_dam_break_trajectory(gpu):
    wl = configs.dam_break(N=16, nsteps=60)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    o = oracle.Oracle.for_workload(wl)
    _compare_run(s, o, 60, 20)


def test_prescribed_boundaries_trajectory(gpu):
    mesh = perturbed_mesh(N=10, seed=4)
    s, o = _pair(mesh, 0, cfl=0.3)
    z, qx, qy = random_state(mesh, seed=9, dz=0.05, q=0.05)
    s.set_state(z, qx, qy)
    o.set_state(z, qx, qy)
    _compare_run(s, o, 40, 10)


def _set_hydrographs(obj, **kw):
    """Hydrographs over a ~30 s run: code 1 an inflow discharge with a
    falling limb (zeta_b at the "use interior" sentinel), code 3 a stage
    series that crosses into the sentinel range (zeta_b < 0) and back, code 2
    a falling stage with its discharge held at the qn > 1000 sentinel.  (A
    ramp from a sentinel to a real discharge would pass through huge
    discharges, so sentinels are held over whole segments.)"""
    t = [0.0, 8.0, 19.0, 54.0]
    obj.set_hydrograph(1, t, [-1000.0] * 4, [-0.02, -0.08, -0.05, -0.01], **kw)
    obj.set_hydrograph(3, t, [0.1, 0.25, -1.0, 0.05], [0.03, 0.05, 0.02, -0.03], **kw)
    obj.set_hydrograph(2, t, [0.3, 0.2, 0.05, 0.1], [2000.0] * 4, **kw)


@pytest.mark.parametrize("wetdry", [0, 1])
def test_hydrograph_trajectory(gpu, wetdry):
    """Time-dependent boundary data (hydrographs on codes 1 and 3) evaluated
    at each step's start, bit for bit against the oracle."""
    mesh = perturbed_mesh(N=10, seed=4)
    s, o = _pair(mesh, wetdry, cfl=0.3)
    _set_hydrographs(s)
    _set_hydrographs(o)
    z, qx, qy = random_state(mesh, seed=9, dz=0.05, q=0.05)
    s.set_state(z, qx, qy)
    o.set_state(z, qx, qy)
    _compare_run(s, o, 40, 10)


def test_forcing_follows_the_handle_time(gpu):
    """Single operators see the hydrograph values and infiltration rate of the
    handle's current time: right after the setters, and after dgr_set_time
    moves a handle back to t = 0 from a later time."""
    mesh = perturbed_mesh(N=10, seed=4)
    z, qx, qy = random_state(mesh, seed=9, dz=0.05, q=0.05)

    def forced(h):
        _set_hydrographs(h)
        h.set_infiltration(1e-5, *GA)

    a = native.Solver(mesh, native.default_params(cfl=0.3))
    forced(a)
    a.set_state(z, qx, qy)
    a.run(5)
    a.set_state(z, qx, qy)
    a.set_time(0.0, native.default_params().dt0)
    b = native.Solver(mesh, native.default_params(cfl=0.3))
    forced(b)
    b.set_state(z, qx, qy)
    c = native.Solver(mesh, native.default_params(cfl=0.3))
    c.set_state(z, qx, qy)
    ra, fa, _ = a.compute2DL()
    rb, fb, _ = b.compute2DL()
    rc, fc, _ = c.compute2DL()
    for x, y in zip(ra + fa, rb + fb):
        assert same(x, y)
    assert not all(same(x, y) for x, y in zip(rb + fb, rc + fc))
    for h in (a, b, c):
        h.close()


def test_hydrograph_rejected_by_adjoint(gpu):
    mesh = perturbed_mesh(N=6, seed=4)
    s = native.Solver(mesh, native.default_params(cfl=0.3))
    _set_hydrographs(s)
    z, qx, qy = random_state(mesh, seed=9, dz=0.05, q=0.05)
    s.set_state(z, qx, qy)
    gel = np.array([0, 5, 9], np.int32)
    obs = np.zeros((1, 2, 3))
    with pytest.raises(native.DgrError) as ei:
        s.adjoint_gauges(4, 2, 2, gel, obs)
    assert ei.value.code == native.DGR_ERR_UNSUPPORTED
    for code in (1, 2, 3):
        s.set_hydrograph(code, [], [], [])
    s.adjoint_gauges(4, 2, 2, gel, obs)


GA = (0.11, 0.43, 0.25)    # suction head psi (m), effective porosity, effective saturation


def test_infiltration_trajectory(gpu):
    """Rain and Green-Ampt infiltration (the step's mean rate computed on the
    device at each step's start) on a wet domain, bit for bit against the
    oracle, with and without the wet/dry scheme."""
    mesh = perturbed_mesh(N=10, seed=4)
    z, qx, qy = random_state(mesh, seed=9, dz=0.05, q=0.05)
    for wetdry in (0, 1):
        s, o = _pair(mesh, wetdry, cfl=0.3, rain=20e-3 / 3600)
        s.set_infiltration(1e-5, *GA)
        o.set_infiltration(1e-5, *GA)
        s.set_state(z, qx, qy)
        o.set_state(z, qx, qy)
        _compare_run(s, o, 40, 20)
        s.close()


def test_infiltration_dries_elements(gpu):
    """Infiltration on the dry-bed dam break: about a third of the elements
    drain dry within 120 steps (none do without it); flags, limiter
    activations and states bit for bit against the oracle."""
    wl = configs.dry_bed_dam_break(N=24, nsteps=120)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_infiltration(1e-6, *GA)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    o = oracle.Oracle.for_workload(wl)
    o.set_infiltration(1e-6, *GA)
    o.set_state(wl.zeta, wl.qx, wl.qy)
    _compare_run(s, o, 120, 40)
    assert 100 < int(np.sum(s.get_wd() == 0)) < 1000


def test_infiltration_failure_matches_oracle(gpu):
    """Infiltration faster than the thin films hold: the oracle stops with a
    wet/dry failure; the device reports the same error code, element and step."""
    wl = configs.dry_bed_dam_break(N=24, nsteps=120)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_infiltration(1e-4, *GA)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    o = oracle.Oracle.for_workload(wl)
    o.set_infiltration(1e-4, *GA)
    o.set_state(wl.zeta, wl.qx, wl.qy)
    for step in range(1, 120):
        try:
            o.run(1)
        except RuntimeError as ex:
            msg = str(ex)
            break
        s.run(1)
        for a, b in zip(s.get_state(), o.get_state()):
            assert same(a, b), "state differs after %d steps" % step
    else:
        pytest.fail("expected the oracle to stop on this input")
    with pytest.raises(native.DgrError) as ei:
        s.run(1)
    code, where = msg.split("error ")[1].split(" at ")
    assert "error %s at %s" % (code, where) in str(ei.value), (msg, str(ei.value))


def test_infiltration_arguments_and_adjoint(gpu):
    mesh = perturbed_mesh(N=6, seed=4)
    s = native.Solver(mesh, native.default_params(cfl=0.3))
    for bad in ((1e-5, 0.0, 0.4, 0.0), (1e-5, 0.1, 1.5, 0.0), (1e-5, 0.1, 0.4, 1.0), (1e-5, 0.1, 0.4, -0.1)):
        with pytest.raises(native.DgrError) as ei:
            s.set_infiltration(*bad)
        assert ei.value.code == native.DGR_ERR_ARG
    z, qx, qy = random_state(mesh, seed=9, dz=0.05, q=0.05)
    s.set_state(z, qx, qy)
    s.set_infiltration(1e-5, *GA)
    gel = np.array([0, 5, 9], np.int32)
    obs = np.zeros((1, 2, 3))
    with pytest.raises(native.DgrError) as ei:
        s.adjoint_gauges(4, 2, 2, gel, obs)
    assert ei.value.code == native.DGR_ERR_UNSUPPORTED
    s.set_infiltration(0.0)
    s.adjoint_gauges(4, 2, 2, gel, obs)
    # off again: the run is the plain one
    s2 = native.Solver(mesh, native.default_params(cfl=0.3))
    for h in (s, s2):
        h.set_state(z, qx, qy)
        h.set_time(0.0, native.default_params().dt0)
        h.run(10)
    for a, b in zip(s.get_state(), s2.get_state()):
        assert same(a, b)


def test_dry_bed_dam_break_trajectory(gpu):
    wl = configs.dry_bed_dam_break(N=24, nsteps=120)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    o = oracle.Oracle.for_workload(wl)
    _compare_run(s, o, 120, 40)


def test_wetdry_failure_matches_oracle(gpu):
    """A sloping beach whose dry films the reference scheme cannot hold: the
    oracle stops on a NaN flux; the device must report the same error code,
    edge and step (the earliest error in the reference's loop order)."""
    mesh = beach_mesh(N=16)
    s, o = _pair(mesh, 1, cfl=0.3)
    z, qx, qy = beach_state(mesh)
    s.set_state(z, qx, qy)
    o.set_state(z, qx, qy)
    for step in range(1, 50):
        try:
            o.run(1)
        except RuntimeError as ex:
            msg = str(ex)
            break
        s.run(1)
        for a, b in zip(s.get_state(), o.get_state()):
            assert same(a, b), "state differs after %d steps" % step
    else:
        pytest.fail("expected the oracle to stop on this input")
    with pytest.raises(native.DgrError) as ei:
        s.run(1)
    code, where = msg.split("error ")[1].split(" at ")
    assert "error %s at %s" % (code, where) in str(ei.value), (msg, str(ei.value))


def test_coastal_tide_rain_trajectory(gpu):
    wl = configs.coastal(N=32, nsteps=60, rain_mm_h=30.0)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    o = oracle.Oracle.for_workload(wl)
    _compare_run(s, o, 60, 20)


def test_dam_break_1000_steps(gpu):
    """The plain dam break (100x100 grid, 20,000 triangles, 1000 SSP-RK2 steps
    at CFL 0.3, wet/dry and limiter on): bitwise equal to the oracle."""
    wl = configs.dam_break(N=100, nsteps=1000)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    o = oracle.Oracle.for_workload(wl, threads=min(16, os.cpu_count() or 1))
    _compare_run(s, o, 1000, 100)
    a = np.concatenate(s.get_state())
    b = np.concatenate(o.get_state())
    assert rel_l2(a, b) == 0.0


def test_one_time_step_host_dt(gpu):
    mesh = perturbed_mesh(N=6, seed=3)
    s, o = _pair(mesh, 0)
    z, qx, qy = random_state(mesh, seed=3, dz=0.05, q=0.05)
    s.set_state(z, qx, qy)
    o.set_state(z, qx, qy)
    for k in range(5):
        s.oneTimeStep2D(0.0, 0.01)
        o.oneTimeStep2D(0.0, 0.01)
    for a, b in zip(s.get_state(), o.get_state()):
        assert same(a, b)


@pytest.mark.parametrize("variant", list(range(native.NUM_VARIANTS)))
def test_stage_variants_bitwise(gpu, variant):
    """Every stage-kernel variant (fused or split flux/RHS kernels, geometry
    recomputed in registers or read from precomputed planes) gives the same
    bits as the oracle: one wet/dry compute2DL and a tide+rain trajectory."""
    mesh = perturbed_mesh(N=8, seed=5)
    s, o = _pair(mesh, 1)
    s.set_variant(variant)
    assert s.variant() == variant
    z, qx, qy = random_state(mesh, seed=15)
    wd = np.random.default_rng(5).integers(0, 2, mesh.num_el).astype(np.int32)
    for x in (s, o):
        x.set_state(z, qx, qy)
        x.set_wd(wd)
    rg, fg, lg = s.compute2DL()
    ro, fo, lo = o.compute2DL()
    for a, b in zip(fg + rg, fo + ro):
        assert same(a, b)
    assert lg == lo
    wl = configs.coastal(N=24, nsteps=40, rain_mm_h=30.0)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_variant(variant)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    _compare_run(s, oracle.Oracle.for_workload(wl), 40, 20)


@pytest.mark.parametrize("threads", [64, 128])
def test_block_sizes_bitwise(gpu, threads):
    """The RHS, node-bound and limiter kernels at 64 or 128 threads per
    workgroup (the 256-element member ranges then span several workgroups)
    give the oracle's bits: a wet/dry dry-bed dam break with PDop acting,
    and a tide+rain run on a mesh whose element count is no multiple of 64."""
    wl = configs.dry_bed_dam_break(N=24, nsteps=60)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_variant(2)
    s.set_block(threads)
    assert s.block() == threads
    s.set_state(wl.zeta, wl.qx, wl.qy)
    _compare_run(s, oracle.Oracle.for_workload(wl), 60, 30)
    wl = configs.coastal(N=23, nsteps=30, rain_mm_h=30.0)
    assert wl.num_el % 64
    for v in range(native.NUM_VARIANTS):
        s = native.Solver(wl.mesh, wl.solver_params())
        s.set_variant(v)
        s.set_block(threads)
        s.set_state(wl.zeta, wl.qx, wl.qy)
        _compare_run(s, oracle.Oracle.for_workload(wl), 30, 30)


def test_autotune_block_keeps_state(gpu):
    """dgr_autotune_block times 64, 128 and 256 threads from the current
    state and restores it; a graph captured before the change is
    recaptured, and the run stays bitwise equal to the oracle."""
    wl = configs.coastal(N=32, nsteps=30)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.set_graph(True)
    o = oracle.Oracle.for_workload(wl)
    _compare_run(s, o, 10, 10)
    chosen, ms = s.autotune_block(3)
    assert chosen in (64, 128, 256) and s.block() == chosen
    assert len(ms) == 3 and all(m > 0 for m in ms), ms
    _compare_run(s, o, 10, 10)
    for t in (64, 128, 256):
        s.set_block(t)
        _compare_run(s, o, 2, 2)
    with pytest.raises(native.DgrError):
        s.set_block(96)


def test_autotune_keeps_state(gpu):
    """dgr_autotune times every variant from the current state and restores
    it: the run that follows is still bitwise equal to the oracle."""
    wl = configs.coastal(N=32, nsteps=30)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    o = oracle.Oracle.for_workload(wl)
    _compare_run(s, o, 10, 10)
    chosen, ms = s.autotune(3)
    assert 0 <= chosen < native.NUM_VARIANTS and s.variant() == chosen
    assert len(ms) == native.NUM_VARIANTS
    assert all(m > 0 for m in ms), ms
    _compare_run(s, o, 20, 10)


@pytest.mark.parametrize("wetdry", [0, 1])
def test_lake_at_rest_on_gpu(gpu, wetdry):
    """Still water over a bumpy bed (walls, 32x32 grid, 500 steps, limiter on):
    the surface stays flat to 1e-13 (the bar on h).  The momentum
    is not exactly balanced by the reference scheme itself: it drifts by
    rounding to ~6e-12 after 500 steps, identically in the oracle's
    reference-math mode (checked here to 1e-11) and bit for bit in its
    device-math mode."""
    wl = configs.lake_at_rest(N=32)
    p = dict(wl.params, wetdry=wetdry)
    s = native.Solver(wl.mesh, native.default_params(**p))
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.run(500)
    z, qx, qy = s.get_state()
    assert np.max(np.abs(z - wl.zeta)) <= 1e-13, np.max(np.abs(z - wl.zeta))
    assert np.max(np.abs(qx)) <= 1e-11 and np.max(np.abs(qy)) <= 1e-11, (np.max(np.abs(qx)), np.max(np.abs(qy)))
    o = oracle.Oracle(wl.mesh, math=oracle.MATH_DEV, **p)
    o.set_state(wl.zeta, wl.qx, wl.qy)
    o.run(500)
    for a, b in zip((z, qx, qy), o.get_state()):
        assert same(a, b)
    r = oracle.Oracle(wl.mesh, math=oracle.MATH_REF, **p)
    r.set_state(wl.zeta, wl.qx, wl.qy)
    r.run(500)
    zr, qxr, qyr = r.get_state()
    assert np.max(np.abs(zr - wl.zeta)) <= 1e-13
    assert np.max(np.abs(qxr)) <= 1e-11 and np.max(np.abs(qyr)) <= 1e-11


def test_set_nodal_steps_like_a_fresh_handle(gpu):
    """A bed and Manning n changed through dgr_set_nodal step exactly as a
    handle created with them: the edge geometry (the bed at each edge's
    midpoint, which every Roe flux reads) follows the new bed.  Coastal
    physics: tide on prescribed edges, wet/dry on, limiter."""
    import dataclasses
    wl = configs.coastal(N=24, nsteps=40)
    rng = np.random.default_rng(11)
    mesh = wl.mesh
    z1 = mesh.z + 0.4 * np.cos(mesh.x / 700.0) * np.sin(mesh.y / 900.0) + 0.05 * rng.standard_normal(mesh.num_nodes)
    n1 = mesh.n * np.exp(0.3 * rng.standard_normal(mesh.num_nodes))
    fresh = native.Solver(dataclasses.replace(mesh, z=z1, n=n1), wl.solver_params())
    moved = native.Solver(mesh, wl.solver_params())
    moved.set_nodal(z=z1, n=n1)
    old = native.Solver(mesh, wl.solver_params())
    outs = []
    for s in (fresh, moved, old):
        s.set_state(wl.zeta, wl.qx, wl.qy)
        rep = s.run(wl.nsteps)
        outs.append((s.get_state(), s.get_wd(), rep.t, rep.dt))
        s.close()
    (a, wa, ta, da), (b, wb, tb, db), (c, _, _, _) = outs
    assert all(same(x, y) for x, y in zip(a, b))
    assert np.array_equal(wa, wb) and ta == tb and da == db
    assert not same(a[0], c[0])   # the new bed changes the run


def test_checkpoint_restart(gpu, tmp_path):
    """A restart file written after 15 steps and read into a new handle gives
    the bits of the uninterrupted run (state, flags, t, dt)."""
    wl = configs.coastal(N=16, rain_mm_h=10.0)
    a = native.Solver(wl.mesh, wl.solver_params())
    a.set_state(wl.zeta, wl.qx, wl.qy)
    a.run(15)
    path = str(tmp_path / "member0.ckpt")
    a.checkpoint_write(path)
    ra = a.run(15)
    b = native.Solver(wl.mesh, wl.solver_params())
    b.checkpoint_read(path)
    rb = b.run(15)
    for x, y in zip(a.get_state(), b.get_state()):
        assert same(x, y)
    assert np.array_equal(a.get_wd(), b.get_wd())
    assert ra.t == rb.t and ra.dt == rb.dt
    with open(path, "r+b") as f:
        f.write(b"XXXXXXXX")
    with pytest.raises(native.DgrError):
        b.checkpoint_read(path)
