"""ChanNet's channel-junction coupling on the device: junction meshes as the
members of an ensemble handle, the junction side of internal_BC
(dgr_couple_channel_ends) against the oracle's restatement (itself bit for
bit the compiled reference's internal_BC, test_oracle_vs_reference.py), and
junctions stepped with a host dt between couplings, as ChanNet's time loop
does (ChanNet/time_evolution.c:552-578), with ChanNet's junction boundaries
(no sentinels) and wet/dry on."""
import numpy as np
import pytest

import oracle
from dgrain import native
from helpers import coupling_network, same

pytestmark = pytest.mark.gpu

PRM = dict(cfl=0.3, wetdry=1)


def _pair(seed):
    meshes, states, spec = coupling_network(seed)
    ens = native.Ensemble(meshes, [native.default_params(bc_sentinels=0, **PRM) for _ in meshes])
    ors = [oracle.Oracle(m, math=oracle.MATH_DEV, bc_sentinels=0, **PRM) for m in meshes]
    for i, (o, st, m) in enumerate(zip(ors, states, meshes)):
        ens.set_state(i, *st)
        o.set_state(*st)
        # ChanNet's junction edges that no channel feeds hold still water:
        # the 2DCode defaults are sentinels, which ChanNet does not read
        z = np.zeros(m.num_edges)
        ens.set_boundary(i, z, z)
        o.set_boundary(z, z)
    return meshes, ens, ors, spec


@pytest.mark.parametrize("seed", range(3))
def test_coupling_matches_oracle(gpu, seed):
    meshes, ens, ors, spec = _pair(seed)
    ga, gq = ens.couple_channel_ends(*spec)
    ga_o, gq_o = oracle.couple_ends(ors, *spec)
    assert same(ga, ga_o) and same(gq, gq_o)
    for i, o in enumerate(ors):
        for a, b in zip(ens.get_boundary(i), o.get_boundary()):
            assert same(a, b)
    ens.close()


def test_coupled_junction_steps(gpu):
    """Ten steps of: couple the junctions to channel ends whose values change
    every step, then one oneTimeStep2D of every junction with the same host
    dt.  States, boundary values and ghost values bit for bit the oracle's."""
    meshes, ens, ors, spec = _pair(7)
    end_off, edge_junc, edge_local, end_dir, end_A, end_Q = spec
    # channel ends in balance with the junctions: the area their edges hold
    # and a small discharge, perturbed every step
    area, _ = oracle.couple_ends(ors, *spec)
    rng = np.random.default_rng(7)
    t, dt = 0.0, 0.05
    for step in range(10):
        A = area * (1 + 0.05 * rng.standard_normal(area.size))
        Q = 0.1 * area * rng.standard_normal(area.size)
        g = ens.couple_channel_ends(end_off, edge_junc, edge_local, end_dir, A, Q)
        go = oracle.couple_ends(ors, end_off, edge_junc, edge_local, end_dir, A, Q)
        for a, b in zip(g, go):
            assert same(a, b), step
        t += dt
        ens.oneTimeStep2D(t, dt)
        for o in ors:
            o.oneTimeStep2D(t, dt)
        for i, o in enumerate(ors):
            for a, b in zip(ens.get_state(i), o.get_state()):
                assert same(a, b), (step, i)
            assert np.array_equal(ens.get_wd(i), o.get_wd()), (step, i)
    ens.close()


def test_coupling_arguments(gpu):
    meshes, ens, ors, spec = _pair(0)
    end_off, edge_junc, edge_local, end_dir, end_A, end_Q = spec
    bad = [(end_off, edge_junc, edge_local, end_dir * 2, end_A, end_Q),
           (end_off, edge_junc + 5, edge_local, end_dir, end_A, end_Q),
           (end_off, edge_junc, edge_local + 100000, end_dir, end_A, end_Q),
           (end_off[::-1].copy(), edge_junc, edge_local, end_dir, end_A, end_Q)]
    for args in bad:
        with pytest.raises(native.DgrError) as ei:
            ens.couple_channel_ends(*args)
        assert ei.value.code == native.DGR_ERR_ARG
    ens.close()
