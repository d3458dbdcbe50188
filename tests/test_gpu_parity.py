"""GPU parity: libdgrain on gfx950, called through the C ABI, against the
oracle in its device-math mode (squares as x*x, H^(4/3) by the shared
deterministic algorithm).  The bar is bit-for-bit equality up to the sign of
zero, for states, RHS, edge fluxes, max_lambda, dt, wet/dry flags and limiter
activations.  (The oracle's reference-math mode is pinned to the compiled
reference in test_oracle_vs_reference.py; the two modes are compared to the
north-star tolerance rel-L2 <= 1e-11 in test_oracle_modes.py.)"""
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
                        rng.uniform(-50, 50, 1000), [0.0, -0.0, 1.0, 8.0, 1e-310, 5e-324, 1e300]])
    y = rng.uniform(-3, 3, x.size)
    out = native.math_selftest(x, y)
    assert same(out[:, 0], np.sqrt(x))                  # IEEE sqrt
    assert same(out[:, 1], x / y)                        # IEEE division
    o43 = np.array([oracle.pow43(v) for v in x])
    assert same(out[:, 2], o43)
    osin = np.array([oracle.dsin(v) for v in x])
    assert same(out[:, 3], osin)


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
    the surface stays flat to 1e-13 (the north-star bar on h).  The momentum
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
