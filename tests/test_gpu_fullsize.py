"""Parity at the benchmark sizes: the headline coastal workload (512x512
cells, 524,288 triangles) bit for bit against the oracle over its first
steps; still water at that size drifts exactly as the oracle's (the
reference scheme's own rounding drift); and the 64-member
ensemble of the batching benchmark (7.3 M triangles) advances every member
bit for bit as its own handle does."""
import os

import numpy as np
import pytest

import oracle
from dgrain import configs, native
from helpers import same

pytestmark = pytest.mark.gpu


def test_coastal_512_flood_bitwise(gpu):
    """The benchmark workload (tide, wet/dry and limiter on, CFL 0.3, flood
    start), 20 steps with dt on the device, every variant's default (the
    autotuned one is one of them): states, wet/dry flags, limiter
    activations, t and dt equal to the oracle's after 10 and 20 steps."""
    wl = configs.coastal(N=512, start="flood")
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    o = oracle.Oracle.for_workload(wl, threads=min(16, os.cpu_count() or 1))
    for done in (10, 20):
        rep = s.run(10)
        t, dt = o.run(10)
        for a, b in zip(s.get_state(), o.get_state()):
            assert same(a, b), "state differs after %d steps" % done
        assert np.array_equal(s.get_wd(), o.get_wd()), done
        for st in (0, 1):
            assert np.array_equal(s.get_modify(st), o.get_modify(st)), done
        assert rep.t == t and rep.dt == dt, (done, rep.t, t, rep.dt, dt)
        print("  %d steps of %d triangles bitwise equal (t=%.6g)" % (done, wl.num_el, t), flush=True)
    s.close()


def test_lake_at_rest_512(gpu):
    """Still water over a bumpy bed with walls at 512x512 cells (L = 10 km),
    50 steps: bit for bit the oracle's.  The reference scheme is well
    balanced only to rounding, and at this size its own drift reaches ~1e-11
    in zeta and ~5e-10 in momentum after 50 steps (the oracle in
    reference-math mode, pinned to the compiled reference, shows the same
    to 2%), so the 1e-13 bar on zeta holds at 32x32 (test_lake_at_rest_on_gpu)
    but not here; both drifts stay below 1e-10 in zeta."""
    wl = configs.lake_at_rest(N=512, L=1.0e4)
    s = native.Solver(wl.mesh, native.default_params(**wl.params))
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.run(50)
    z, qx, qy = s.get_state()
    threads = min(16, os.cpu_count() or 1)
    o = oracle.Oracle(wl.mesh, math=oracle.MATH_DEV, threads=threads, **wl.params)
    o.set_state(wl.zeta, wl.qx, wl.qy)
    o.run(50)
    for a, b in zip((z, qx, qy), o.get_state()):
        assert same(a, b)
    r = oracle.Oracle(wl.mesh, math=oracle.MATH_REF, threads=threads, **wl.params)
    r.set_state(wl.zeta, wl.qx, wl.qy)
    r.run(50)
    dz, dz_ref = np.max(np.abs(z - wl.zeta)), np.max(np.abs(r.get_state()[0] - wl.zeta))
    assert dz <= 1e-10 and abs(dz - dz_ref) <= 0.02 * dz_ref, (dz, dz_ref)
    print("  zeta drift %.3e (reference math %.3e)" % (dz, dz_ref))
    s.close()


def test_benchmark_ensemble_members_match_solo(gpu):
    """The 64 members of the batching benchmark (N log-uniform in [64, 512])
    under graph replay, 4 steps: every member's state, flags, t and dt equal
    its solo run's."""
    wls = configs.ensemble(n=64, seed=3)
    ens = native.Ensemble([w.mesh for w in wls], [w.solver_params() for w in wls])
    for i, w in enumerate(wls):
        ens.set_state(i, w.zeta, w.qx, w.qy)
    ens.set_graph(True)
    reps = ens.run(4)
    for i, w in enumerate(wls):
        s = native.Solver(w.mesh, w.solver_params())
        s.set_variant(ens.variant())
        s.set_state(w.zeta, w.qx, w.qy)
        r = s.run(4)
        for a, b in zip(ens.get_state(i), s.get_state()):
            assert same(a, b), "member %d (%s) differs from its solo run" % (i, w.name)
        assert np.array_equal(ens.get_wd(i), s.get_wd()), i
        assert reps[i].t == r.t and reps[i].dt == r.dt, i
        s.close()
    print("  %d members, %d triangles, bitwise equal to their solo runs" % (len(wls), sum(w.num_el for w in wls)))
    ens.close()
