"""GPU parity on unstructured meshes: Delaunay triangulations of random
nodes with shuffled node and element numbering, node valences from 1 to
~10 and every boundary code (the structured test meshes have valences of at
most 8 and ordered numbering).  Forward results bit for bit against the
oracle in device-math mode; ensemble members bit for bit against their solo
runs; the adjoint against central differences of the GPU forward."""
import numpy as np
import pytest

import oracle
from dgrain import configs, native
from helpers import delaunay_mesh, random_state, same

pytestmark = pytest.mark.gpu


def _workload(seed, nsteps=40, wetdry=1, n_in=80):
    """Tide on the code-2 edges, rain, wet/dry and limiter, a generic start
    (still water plus small random elevations and flows)."""
    mesh = delaunay_mesh(seed=seed, n_in=n_in)
    rng = np.random.default_rng(seed + 1)
    n3 = 3 * mesh.num_el
    zeta = 0.05 * rng.standard_normal(n3)
    qx, qy = 0.05 * rng.standard_normal(n3), 0.05 * rng.standard_normal(n3)
    params = dict(cfl=0.3, wetdry=wetdry, tide_code=2, tide_datum=0.0, tide_amp=0.3, tide_period=600.0,
                  tide_phase=0.0, rain=30.0 / 1000.0 / 3600.0)
    return configs.Workload("delaunay_%d" % seed, mesh, zeta, qx, qy, params, nsteps)


@pytest.mark.parametrize("wetdry", [0, 1])
@pytest.mark.parametrize("variant", list(range(native.NUM_VARIANTS)))
def test_unstructured_compute2DL(gpu, variant, wetdry):
    """compute2DL (edge fluxes, RHS, max |lambda|) of every kernel variant
    bit for bit, with random wet/dry flags when wet/dry is on."""
    for seed in (0, 1):
        mesh = delaunay_mesh(seed=seed)
        s = native.Solver(mesh, native.default_params(wetdry=wetdry))
        o = oracle.Oracle(mesh, math=oracle.MATH_DEV, wetdry=wetdry)
        s.set_variant(variant)
        st = random_state(mesh, seed=seed + 20, dz=0.3, q=0.3)
        for x in (s, o):
            x.set_state(*st)
        if wetdry:
            wd = np.random.default_rng(seed).integers(0, 2, mesh.num_el).astype(np.int32)
            s.set_wd(wd)
            o.set_wd(wd)
        rg, fg, lg = s.compute2DL()
        ro, fo, lo = o.compute2DL()
        for a, b in zip(fg + rg, fo + ro):
            assert same(a, b)
        assert lg == lo
        s.close()


@pytest.mark.parametrize("variant", list(range(native.NUM_VARIANTS)))
def test_unstructured_trajectory(gpu, variant):
    """60 steps with tide, rain, wet/dry and limiter, dt on the device:
    states, flags, limiter activations, t and dt equal to the oracle's."""
    wl = _workload(2, nsteps=60)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_variant(variant)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    o = oracle.Oracle.for_workload(wl)
    for done in (20, 40, 60):
        rep = s.run(20)
        t, dt = o.run(20)
        for a, b in zip(s.get_state(), o.get_state()):
            assert same(a, b), "state differs after %d steps" % done
        assert np.array_equal(s.get_wd(), o.get_wd()), done
        for st in (0, 1):
            assert np.array_equal(s.get_modify(st), o.get_modify(st)), done
        assert rep.t == t and rep.dt == dt, (done, rep.t, t, rep.dt, dt)
    s.close()


def test_unstructured_ensemble_members_match_solo(gpu):
    """Three Delaunay members of different sizes and a structured one in one
    ensemble under graph replay, 30 steps: each member's state, flags, t and
    dt equal its solo run's."""
    wls = [_workload(3, n_in=60), configs.coastal(N=12, nsteps=30), _workload(4, n_in=300), _workload(5, n_in=30)]
    ens = native.Ensemble([w.mesh for w in wls], [w.solver_params() for w in wls])
    for i, w in enumerate(wls):
        ens.set_state(i, w.zeta, w.qx, w.qy)
    ens.set_graph(True)
    reps = ens.run(30)
    for i, w in enumerate(wls):
        s = native.Solver(w.mesh, w.solver_params())
        s.set_variant(ens.variant())
        s.set_state(w.zeta, w.qx, w.qy)
        r = s.run(30)
        for a, b in zip(ens.get_state(i), s.get_state()):
            assert same(a, b), "member %d differs from its solo run" % i
        assert np.array_equal(ens.get_wd(i), s.get_wd()), i
        assert reps[i].t == r.t and reps[i].dt == r.dt, i
        s.close()
    ens.close()


@pytest.mark.parametrize("param", ["z", "n"])
def test_unstructured_adjoint_directional_derivatives(gpu, param):
    """The gauge-misfit gradient on a Delaunay mesh against central
    differences of the GPU forward, 3 smooth random directions; tolerance
    relative 1e-6 for z (eps 1e-6 of the field) and 2e-6 for n (eps 1e-4:
    over 30 steps the n gradient is ~1e-11, and the central differences at
    eps 1e-4 and 1e-3 both scatter by ~1e-6 around it).  Resident tapes
    and recomputed segments give the same bits."""
    nsteps, every = 30, 5
    wl = _workload(6, nsteps=nsteps)
    rng = np.random.default_rng(7)
    gauges = np.sort(rng.choice(wl.num_el, 8, replace=False)).astype(np.int32)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_nodal(n=wl.mesh.n * np.exp(0.2 * rng.standard_normal(wl.mesh.num_nodes)))
    s.set_state(wl.zeta, wl.qx, wl.qy)
    obs = s.gauge_run(nsteps, every, gauges)
    s.close()

    def loss_at(**kw):
        s = native.Solver(wl.mesh, wl.solver_params())
        s.set_nodal(**kw)
        s.set_state(wl.zeta, wl.qx, wl.qy)
        h = s.gauge_run(nsteps, every, gauges)
        s.close()
        return float(((h - obs) ** 2).sum())

    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    loss, gz, gn = s.adjoint_gauges(nsteps, every, 10, gauges, obs)
    s.set_tape_budget(0)
    loss_r, gz_r, gn_r = s.adjoint_gauges(nsteps, every, 10, gauges, obs)
    s.close()
    assert loss_r[0] == loss[0] and np.array_equal(gz_r, gz) and np.array_equal(gn_r, gn)
    g = gz if param == "z" else gn
    base = wl.mesh.z if param == "z" else wl.mesh.n
    x = (wl.mesh.x - wl.mesh.x.min()) / np.ptp(wl.mesh.x)
    y = (wl.mesh.y - wl.mesh.y.min()) / np.ptp(wl.mesh.y)
    worst = 0.0
    for _ in range(3):
        kx, ky = rng.integers(0, 3, 2)
        d = np.cos(np.pi * kx * x + rng.uniform(0, 6.3)) * np.cos(np.pi * ky * y + rng.uniform(0, 6.3))
        d *= np.abs(base).mean()
        eps = 1e-6 if param == "z" else 1e-4
        fd = (loss_at(**{param: base + eps * d}) - loss_at(**{param: base - eps * d})) / (2 * eps)
        ad = float(g @ d)
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        print("  %s: adjoint %.12e  central difference %.12e  rel %.2e" % (param, ad, fd, rel), flush=True)
        worst = max(worst, rel)
    assert worst <= (1e-6 if param == "z" else 2e-6), worst


@pytest.mark.parametrize("variant", list(range(native.NUM_VARIANTS)))
def test_renumbered_mesh_runs_the_same(gpu, variant):
    """The coastal workload (32x32 cells) with node and element numbers
    shuffled, and the same mesh renumbered by locality_order: 40 steps of
    each on the GPU are bit for bit the same run (states, flags, t, dt),
    and the oracle's on the shuffled mesh."""
    from dgrain import mesh as M
    wl = configs.coastal(N=32, nsteps=40, rain_mm_h=20.0)
    rng = np.random.default_rng(8)
    m0 = wl.mesh
    pn = rng.permutation(m0.num_nodes)
    inv = np.empty_like(pn)
    inv[pn] = np.arange(pn.size)
    pe = rng.permutation(m0.num_el)
    lists = [inv[np.asarray(a)] if np.size(a) else None for a in (m0.in1, m0.in2, m0.out)]
    sh = M.from_elements(m0.x[pn], m0.y[pn], m0.z[pn], m0.n[pn], inv[m0.elto_vert.reshape(-1, 3)[pe]], *lists)
    st = [np.asarray(a).reshape(-1, 3)[pe].ravel() for a in (wl.zeta, wl.qx, wl.qy)]
    m, rn = M.locality_order(sh)
    runs = []
    for mesh, s0 in ((sh, st), (m, [rn.midpoints_to_new(a) for a in st])):
        s = native.Solver(mesh, wl.solver_params())
        s.set_variant(variant)
        s.set_state(*s0)
        rep = s.run(40)
        runs.append((s.get_state(), s.get_wd(), rep.t, rep.dt))
        s.close()
    (sa, wa, ta, da), (sb, wb, tb, db) = runs
    for x, y in zip(sa, sb):
        assert same(x, rn.midpoints_to_old(y))
    assert np.array_equal(wa, rn.elements_to_old(wb)) and ta == tb and da == db
    o = oracle.Oracle(sh, math=oracle.MATH_DEV, **wl.params)
    o.set_state(*st)
    t, dt = o.run(40)
    for x, y in zip(sa, o.get_state()):
        assert same(x, y)
    assert ta == t and da == dt


@pytest.mark.parametrize("variant", list(range(native.NUM_VARIANTS)))
def test_renumbered_edges_wet_run_the_same(gpu, variant):
    """locality_order(edges="locality") on a wet Delaunay workload (tide,
    rain, no both-dry edge): 40 GPU steps on the renumbered mesh are the
    oracle's run of the original mesh, renumbered, bit for bit."""
    from dgrain import mesh as M
    wl = _workload(9, n_in=400)
    m, rn = M.locality_order(wl.mesh, edges="locality")
    s = native.Solver(m, wl.solver_params())
    s.set_variant(variant)
    s.set_state(*[rn.midpoints_to_new(a) for a in (wl.zeta, wl.qx, wl.qy)])
    rep = s.run(40)
    o = oracle.Oracle.for_workload(wl)
    t, dt = o.run(40)
    for x, y in zip(s.get_state(), o.get_state()):
        assert same(rn.midpoints_to_old(x), y)
    assert np.array_equal(rn.elements_to_old(s.get_wd()), o.get_wd())
    assert rep.t == t and rep.dt == dt
    s.close()


def _fan_mesh(K):
    """A centre node shared by K triangles inside a ring of 2K annulus
    triangles."""
    from dgrain import mesh as M
    ang = 2 * np.pi * np.arange(K) / K
    x = np.concatenate([[0.0], 10 * np.cos(ang), 20 * np.cos(ang + np.pi / K)])
    y = np.concatenate([[0.0], 10 * np.sin(ang), 20 * np.sin(ang + np.pi / K)])
    ev = []
    for i in range(K):
        a, b = 1 + i, 1 + (i + 1) % K
        c, d = 1 + K + i, 1 + K + (i + 1) % K
        ev += [(0, a, b), (a, c, b), (b, c, d)]
    rng = np.random.default_rng(K)
    z = 2.0 + rng.uniform(-0.3, 0.3, x.size)
    n = rng.uniform(0.01, 0.04, x.size)
    return M.from_elements(x, y, z, n, np.array(ev, np.int32))


@pytest.mark.parametrize("K", [7, 12, 24])
def test_node_valence_paths(gpu, K):
    """Nodes shared by 7, 12 and 24 elements (the node bounds walk each
    node's element list in order): 40 wet/dry steps of each, bit for bit the
    oracle's."""
    mesh = _fan_mesh(K)
    rng = np.random.default_rng(K + 1)
    n3 = 3 * mesh.num_el
    wl = configs.Workload("fan_%d" % K, mesh, 0.1 * rng.standard_normal(n3), 0.2 * rng.standard_normal(n3),
                          0.2 * rng.standard_normal(n3), dict(cfl=0.3, wetdry=1), 40)
    for variant in (0, 2):
        s = native.Solver(wl.mesh, wl.solver_params())
        s.set_variant(variant)
        s.set_state(wl.zeta, wl.qx, wl.qy)
        o = oracle.Oracle.for_workload(wl)
        rep = s.run(40)
        t, dt = o.run(40)
        for a, b in zip(s.get_state(), o.get_state()):
            assert same(a, b)
        assert np.array_equal(s.get_wd(), o.get_wd())
        for st in (0, 1):
            assert np.array_equal(s.get_modify(st), o.get_modify(st))
        assert rep.t == t and rep.dt == dt
        s.close()
