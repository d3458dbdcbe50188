"""Channel coupling on the device.

ChanNet's channel-junction coupling: junction meshes as the members of an
ensemble handle, the junction side of internal_BC (dgr_couple_channel_ends)
against the oracle's restatement (itself bit for bit the compiled
reference's internal_BC, test_oracle_vs_reference.py), and junctions stepped
with a host dt between couplings, as ChanNet's time loop does
(ChanNet/time_evolution.c:552-578), with ChanNet's junction boundaries (no
sentinels) and wet/dry on.

DG-SWEM's coupling edge with DGSHED's channels (dgr_couple_dgshed) against
the oracle's restatement; that restatement is pinned by closed forms only
(test_oracle_closed_forms.py: the Fortran side and DGSHED's headers are
absent here)."""
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


# ---------------------------------------------------------------------------
# DG-SWEM's coupling edge with DGSHED's 1-D channels (src/couple2dgshed.F,
# DGSHED/couple_domains.c:248-335), with the members of a handle where
# DG-SWEM's mesh stands: the coupled edges are prescribed boundary edges
# (no sentinels), their boundary data come from the channel ends, and the
# channel ends' ghost values from the members' edge traces.
# ---------------------------------------------------------------------------
def dgshed_setup(seed, nmem=2):
    from helpers import perturbed_mesh, random_state
    rng = np.random.default_rng(seed)
    meshes = [perturbed_mesh(N=5 + i, seed=seed + i) for i in range(nmem)]
    states = [random_state(m, seed=500 + seed + i, dz=0.3, q=0.3) for i, m in enumerate(meshes)]
    mem, loc = [], []
    for i, m in enumerate(meshes):
        presc = np.nonzero(np.asarray(m.bdry) > 0)[0]
        for k in rng.choice(presc, 3, replace=False):
            mem.append(i)
            loc.append(int(k))
    n = len(mem)
    conn = rng.integers(0, 2, n).astype(np.int32)
    b = rng.uniform(1.0, 5.0, n)
    m1 = np.where(rng.random(n) < 0.4, 0.0, rng.uniform(0.2, 2.0, n))
    m2 = np.where(m1 == 0, 0.0, rng.uniform(0.2, 2.0, n))
    h = rng.uniform(0.5, 3.0, n)
    zc = h - rng.uniform(-0.2, 0.2, n)   # channel surfaces near the datum, as the 2-D surface
    A = h * b + 0.5 * (m1 + m2) * h * h
    chan = np.stack([A, rng.uniform(-0.5, 0.5, n), b, m1, m2, zc], axis=1)
    geom = np.stack([b * rng.uniform(0.9, 1.1, n), m1, m2, zc + rng.uniform(-0.1, 0.1, n)], axis=1)
    return meshes, states, np.array(mem, np.int32), np.array(loc, np.int32), conn, chan, geom


def _dgshed_pair(seed):
    meshes, states, mem, loc, conn, chan, geom = dgshed_setup(seed)
    ens = native.Ensemble(meshes, [native.default_params(bc_sentinels=0, **PRM) for _ in meshes])
    ors = [oracle.Oracle(m, math=oracle.MATH_DEV, bc_sentinels=0, **PRM) for m in meshes]
    for i, (o, st, m) in enumerate(zip(ors, states, meshes)):
        ens.set_state(i, *st)
        o.set_state(*st)
        # prescribed edges that no channel feeds hold still water (the 2DCode
        # defaults are sentinels, read as values without sentinels)
        z = np.zeros(m.num_edges)
        ens.set_boundary(i, z, z)
        o.set_boundary(z, z)
    return ens, ors, mem, loc, conn, chan, geom


def _dgshed_oracle(ors, mem, loc, conn, chan, geom):
    swe, ghost = np.zeros((mem.size, 2)), np.zeros((mem.size, 2))
    for i, o in enumerate(ors):
        sel = mem == i
        if sel.any():
            swe[sel], ghost[sel] = oracle.couple_dgshed(o, loc[sel], conn[sel], chan[sel], geom[sel])
    return swe, ghost


@pytest.mark.parametrize("seed", range(3))
def test_dgshed_coupling_matches_oracle(gpu, seed):
    """The exchange of one step (channel values into the coupled edges'
    boundary data, edge traces into the channel ends' ghost values), two
    members, rectangular and trapezoidal channels, both end types: bit for
    bit the oracle's."""
    ens, ors, mem, loc, conn, chan, geom = _dgshed_pair(seed)
    swe, ghost = ens.couple_dgshed(mem, loc, conn, chan, geom)
    swe_o, ghost_o = _dgshed_oracle(ors, mem, loc, conn, chan, geom)
    assert same(swe, swe_o) and same(ghost, ghost_o)
    for i, o in enumerate(ors):
        for a, b in zip(ens.get_boundary(i), o.get_boundary()):
            assert same(a, b)
    ens.close()


def test_dgshed_coupled_steps(gpu):
    """Ten steps of DG-SWEM's order (src/DG_timestep.F:34-59): the 2-D step
    with the current channel values as boundary data, then the exchange;
    the channel ends take the ghost values back (a host stand-in for the
    1-D channel step that moves A and Q towards them).  States, flags and
    exchanged values bit for bit the oracle's."""
    ens, ors, mem, loc, conn, chan, geom = _dgshed_pair(11)
    chan = chan.copy()
    ens.couple_dgshed(mem, loc, conn, chan, geom)
    _dgshed_oracle(ors, mem, loc, conn, chan, geom)
    t, dt = 0.0, 0.02
    for step in range(10):
        t += dt
        ens.oneTimeStep2D(t, dt)
        for o in ors:
            o.oneTimeStep2D(t, dt)
        for i, o in enumerate(ors):
            for a, b in zip(ens.get_state(i), o.get_state()):
                assert same(a, b), (step, i)
            assert np.array_equal(ens.get_wd(i), o.get_wd()), (step, i)
        swe, ghost = ens.couple_dgshed(mem, loc, conn, chan, geom)
        swe_o, ghost_o = _dgshed_oracle(ors, mem, loc, conn, chan, geom)
        assert same(swe, swe_o) and same(ghost, ghost_o), step
        chan[:, 0] = 0.8 * chan[:, 0] + 0.2 * ghost[:, 0]
        chan[:, 1] = 0.8 * chan[:, 1] + 0.2 * np.where(conn == 0, ghost[:, 1], -ghost[:, 1])
    ens.close()


def test_dgshed_coupling_arguments(gpu):
    """Edges that are not boundary edges, members or edges out of range and
    end types other than 0/1 are refused; a refused edge keeps its data."""
    from helpers import perturbed_mesh
    ens, ors, mem, loc, conn, chan, geom = _dgshed_pair(0)
    m0 = perturbed_mesh(N=5, seed=0)
    inner = int(np.nonzero(np.asarray(m0.edg_to_els[0::2]) != np.asarray(m0.edg_to_els[1::2]))[0][0])
    before = ens.get_boundary(0)
    for args in [(mem[:1], np.array([inner], np.int32), conn[:1], chan[:1], geom[:1]),
                 (mem[:1] + 5, loc[:1], conn[:1], chan[:1], geom[:1]),
                 (mem[:1], loc[:1] + 10 ** 6, conn[:1], chan[:1], geom[:1]),
                 (mem[:1], loc[:1], conn[:1] + 2, chan[:1], geom[:1])]:
        with pytest.raises(native.DgrError):
            ens.couple_dgshed(*args)
    for a, b in zip(ens.get_boundary(0), before):
        assert same(a, b)
    ens.close()
