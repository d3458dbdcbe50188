"""Ensembles and captured steps on the GPU.

An ensemble handle advances members of uneven sizes, each with its own
forcing, time and dt, with one launch per kernel for all of them.  The bar:
every member is bit-for-bit its own single-instance run (state, WD flags,
limiter activations, t, dt), and therefore the oracle's.  A step replayed
from a captured hipGraph gives the same bits as the enqueued launches."""
import numpy as np
import pytest

import oracle
from dgrain import configs, native
from helpers import same

pytestmark = pytest.mark.gpu


def _members():
    # uneven sizes (some not multiples of the 256-element workgroup), own
    # tide amplitude, rain, bumps and n field per member, one dam break
    wls = configs.ensemble(n=5, seed=11, nmin=6, nmax=40, nsteps=30)
    wls.append(configs.dam_break(N=13, nsteps=30))
    return wls


@pytest.mark.parametrize("variant", list(range(native.NUM_VARIANTS)))
def test_ensemble_members_match_solo(gpu, variant):
    wls = _members()
    ens = native.Ensemble([w.mesh for w in wls], [w.solver_params() for w in wls])
    ens.set_variant(variant)
    for i, w in enumerate(wls):
        ens.set_state(i, w.zeta, w.qx, w.qy)
    reps = ens.run(15)
    reps = ens.run(15)
    for i, w in enumerate(wls):
        s = native.Solver(w.mesh, w.solver_params())
        s.set_variant(variant)
        s.set_state(w.zeta, w.qx, w.qy)
        r = s.run(30)
        for a, b in zip(ens.get_state(i), s.get_state()):
            assert same(a, b), "member %d (%s) state differs from its solo run" % (i, w.name)
        assert np.array_equal(ens.get_wd(i), s.get_wd()), i
        for st in (0, 1):
            assert np.array_equal(ens.get_modify(i, st), s.get_modify(st)), i
        assert reps[i].t == r.t and reps[i].dt == r.dt, (i, reps[i].t, r.t, reps[i].dt, r.dt)
        s.close()
    print("  %d members, %d elements, bitwise equal to their solo runs" % (len(wls), sum(w.num_el for w in wls)))


@pytest.mark.parametrize("threads", [64, 128])
def test_ensemble_block_sizes_match_solo(gpu, threads):
    """Members of an ensemble stepped with 64- or 128-thread workgroups in the
    RHS, node-bound and limiter kernels (several workgroups per 256-element
    member range, padding workgroups at a member's end) are bit for bit their
    solo runs at the default 256, with and without the captured step."""
    wls = _members()
    for graph in (False, True):
        ens = native.Ensemble([w.mesh for w in wls], [w.solver_params() for w in wls])
        ens.set_variant(2)
        ens.set_block(threads)
        ens.set_graph(graph)
        for i, w in enumerate(wls):
            ens.set_state(i, w.zeta, w.qx, w.qy)
        reps = ens.run(30)
        for i, w in enumerate(wls):
            s = native.Solver(w.mesh, w.solver_params())
            s.set_variant(2)
            s.set_state(w.zeta, w.qx, w.qy)
            r = s.run(30)
            for a, b in zip(ens.get_state(i), s.get_state()):
                assert same(a, b), "member %d state differs at %d threads" % (i, threads)
            assert np.array_equal(ens.get_wd(i), s.get_wd()), i
            assert reps[i].t == r.t and reps[i].dt == r.dt, i
            s.close()
        ens.close()


def test_ensemble_member_hydrographs(gpu):
    """Per-member hydrographs (none on member 0, different ones on members 1
    and 2) under graph replay: each member bit for bit its solo run."""
    from helpers import perturbed_mesh, random_state
    meshes = [perturbed_mesh(N=n, seed=sd) for n, sd in ((6, 1), (9, 2), (7, 3))]
    params = [native.default_params(cfl=0.3) for _ in meshes]
    hyd = {1: (1, [0.0, 6.0, 30.0], [-1000.0] * 3, [-0.02, -0.09, 0.0]),
           2: (3, [0.0, 4.0, 9.0, 40.0], [0.05, 0.2, -1.0, 0.1], [0.02, 0.04, 0.01, -0.02])}
    states = [random_state(m, seed=20 + i, dz=0.05, q=0.05) for i, m in enumerate(meshes)]
    ens = native.Ensemble(meshes, params)
    for i, (code, t, zb, qn) in hyd.items():
        ens.set_hydrograph(code, t, zb, qn, member=i)
    for i, st in enumerate(states):
        ens.set_state(i, *st)
    ens.set_graph(True)
    reps = ens.run(12)
    reps = ens.run(13)
    for i, m in enumerate(meshes):
        s = native.Solver(m, params[i])
        if i in hyd:
            code, t, zb, qn = hyd[i]
            s.set_hydrograph(code, t, zb, qn)
        s.set_state(*states[i])
        r = s.run(25)
        for a, b in zip(ens.get_state(i), s.get_state()):
            assert same(a, b), "member %d differs from its solo run" % i
        assert reps[i].t == r.t and reps[i].dt == r.dt, i
        s.close()


def test_ensemble_member_infiltration(gpu):
    """Per-member Green-Ampt parameters (none on member 1): each member bit
    for bit its solo run under graph replay."""
    from helpers import perturbed_mesh, random_state
    meshes = [perturbed_mesh(N=n, seed=sd) for n, sd in ((6, 1), (9, 2), (7, 3))]
    params = [native.default_params(cfl=0.3, rain=10e-3 / 3600) for _ in meshes]
    ga = {0: (1e-5, 0.11, 0.43, 0.25), 2: (4e-6, 0.2, 0.35, 0.1)}
    states = [random_state(m, seed=20 + i, dz=0.05, q=0.05) for i, m in enumerate(meshes)]
    ens = native.Ensemble(meshes, params)
    for i, g in ga.items():
        ens.set_infiltration(*g, member=i)
    for i, st in enumerate(states):
        ens.set_state(i, *st)
    ens.set_graph(True)
    reps = ens.run(12)
    reps = ens.run(13)
    for i, m in enumerate(meshes):
        s = native.Solver(m, params[i])
        if i in ga:
            s.set_infiltration(*ga[i])
        s.set_state(*states[i])
        r = s.run(25)
        for a, b in zip(ens.get_state(i), s.get_state()):
            assert same(a, b), "member %d differs from its solo run" % i
        assert reps[i].t == r.t and reps[i].dt == r.dt, i
        s.close()


def test_ensemble_matches_oracle(gpu):
    wls = _members()[:3]
    ens = native.Ensemble([w.mesh for w in wls], [w.solver_params() for w in wls])
    for i, w in enumerate(wls):
        ens.set_state(i, w.zeta, w.qx, w.qy)
    ens.set_graph(True)
    for chunk in range(3):
        reps = ens.run(10)
    for i, w in enumerate(wls):
        o = oracle.Oracle.for_workload(w)
        t, dt = o.run(30)
        for a, b in zip(ens.get_state(i), o.get_state()):
            assert same(a, b), i
        assert np.array_equal(ens.get_wd(i), o.get_wd()), i
        assert reps[i].t == t and reps[i].dt == dt


def test_graph_replays_match_oracle(gpu):
    """Replays of the captured step, with eager work (state reads, a
    standalone compute2DL) between them, checked against the oracle after
    every replay chunk."""
    wl = configs.coastal(N=20, nsteps=40, rain_mm_h=20.0)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.set_graph(True)
    o = oracle.Oracle.for_workload(wl)
    done = 0
    for k in (1, 5, 10, 24):
        rep = s.run(k)
        t, dt = o.run(k)
        done += k
        for a, b in zip(s.get_state(), o.get_state()):
            assert same(a, b), "state differs after %d graph steps" % done
        assert rep.t == t and rep.dt == dt
        s.compute2DL()   # eager launches between replays
    for v in range(native.NUM_VARIANTS):   # a variant change recaptures
        s.set_variant(v)
        rep = s.run(2)
        t, dt = o.run(2)
        for a, b in zip(s.get_state(), o.get_state()):
            assert same(a, b), "state differs after variant %d replays" % v


def test_ensemble_member_error_is_isolated(gpu):
    """A member that stops on a wet/dry failure reports it (with the edge the
    oracle names); the other members keep their solo trajectories."""
    from helpers import beach_mesh, beach_state
    good = configs.coastal(N=12, nsteps=20)
    bmesh = beach_mesh(N=16)
    bp = native.default_params(wetdry=1, cfl=0.3)
    ens = native.Ensemble([good.mesh, bmesh], [good.solver_params(), bp])
    ens.set_state(0, good.zeta, good.qx, good.qy)
    bz, bqx, bqy = beach_state(bmesh)
    ens.set_state(1, bz, bqx, bqy)
    o = oracle.Oracle(bmesh, math=oracle.MATH_DEV, wetdry=1, cfl=0.3)
    o.set_state(bz, bqx, bqy)
    msg = None
    for step in range(1, 50):
        try:
            o.run(1)
        except RuntimeError as ex:
            msg = str(ex)
            break
    assert msg is not None
    with pytest.raises(native.DgrError):
        ens.run(step)
    code, where = msg.split("error ")[1].split(" at ")
    with pytest.raises(native.DgrError) as ei:
        ens.report(1)
    assert "error %s at %s" % (code, where) in str(ei.value), (msg, str(ei.value))
    rep0 = ens.report(0)
    s = native.Solver(good.mesh, good.solver_params())
    s.set_state(good.zeta, good.qx, good.qy)
    r = s.run(step)
    for a, b in zip(ens.get_state(0), s.get_state()):
        assert same(a, b)
    assert rep0.t == r.t
