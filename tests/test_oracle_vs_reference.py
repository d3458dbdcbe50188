"""The oracle against the compiled reference operators themselves
(oracle/_ref, built from the reference sources by oracle/Makefile).  Random
inputs beyond the golden vectors; bit-for-bit in reference-math mode.
Skipped where the reference sources are not mounted."""
import numpy as np
import pytest

import oracle
from helpers import coupling_network, delaunay_mesh, perturbed_mesh, random_state, same
from conftest import require_ref

pytestmark = pytest.mark.ref


@pytest.fixture(autouse=True)
def _need_ref():
    require_ref()


@pytest.mark.parametrize("seed", range(4))
def test_compute2DL_and_limiter_random(seed):
    m = perturbed_mesh(N=7 + seed, seed=seed)
    ref = oracle.Ref2D(m)
    o = oracle.Oracle(m, math=oracle.MATH_REF)
    st = random_state(m, seed=100 + seed)
    ref.set_state(*st)
    o.set_state(*st)
    r1, f1, l1 = ref.compute2DL()
    r2, f2, l2 = o.compute2DL()
    assert all(same(a, b) for a, b in zip(f1, f2))
    assert all(same(a, b) for a, b in zip(r1, r2))
    assert l1 == l2
    ref.SlopeLimiter()
    o.SlopeLimiter()
    assert all(same(a, b) for a, b in zip(ref.get_state(), o.get_state()))


def test_roe_flux_random():
    m = perturbed_mesh(N=3)
    oracle.Ref2D(m)
    rng = np.random.default_rng(5)
    for _ in range(2000):
        a = list(rng.uniform(-1, 1, 6)) + [rng.uniform(1.2, 4)] + [0.6, 0.8, rng.choice([0.0, 9.81])]
        F1, l1 = oracle.ref_roe_flux(*a)
        F2, l2 = oracle.roe_flux(*a)
        assert same(F1, F2) and l1 == l2


def test_2dcode_steps_random():
    m = perturbed_mesh(N=9, seed=8)
    ref = oracle.Ref2D(m)
    o = oracle.Oracle(m, math=oracle.MATH_REF)
    st = random_state(m, seed=3, dz=0.05, q=0.05)
    ref.set_state(*st)
    o.set_state(*st)
    for _ in range(15):
        ref.oneTimeStep2D(0.0, 0.05)
        o.oneTimeStep2D(0.0, 0.05)
    assert all(same(a, b) for a, b in zip(ref.get_state(), o.get_state()))


def test_wetdry_operators_against_both_references():
    """wetDryStatus2D and PDop2D of 2DCode and ChanNet agree with the oracle."""
    m = perturbed_mesh(N=8, seed=6, depth=0.3)
    rng = np.random.default_rng(6)
    st = list(random_state(m, seed=6, dz=0.25, q=0.05))
    wd = rng.integers(-1, 2, m.num_el).astype(np.int32)
    outs = []
    for impl in (oracle.Ref2D(m), oracle.RefChan(m), oracle.Oracle(m, math=oracle.MATH_REF, wetdry=1)):
        impl.set_state(*st)
        impl.set_wd(wd)
        impl.wetDryStatus2D()
        impl.PDop2D()
        outs.append((impl.get_wd(), impl.get_state()))
    for wd_i, s_i in outs[1:]:
        assert np.array_equal(wd_i, outs[0][0])
        assert all(same(a, b) for a, b in zip(s_i, outs[0][1]))


def test_wetdry_flux_against_channet():
    """The -DWDON flux branches (compute2DL.c:201-293) on walls-only meshes with
    mixed wet/dry flags, against ChanNet's compiled compute2DL."""
    for seed in range(3):
        m = perturbed_mesh(N=8, seed=seed, codes=False)
        st = random_state(m, seed=seed + 40, dz=0.2, q=0.3)
        wd = np.random.default_rng(seed).integers(0, 2, m.num_el).astype(np.int32)
        ch = oracle.RefChan(m)
        o = oracle.Oracle(m, math=oracle.MATH_REF, wetdry=1)
        for impl in (ch, o):
            impl.set_state(*st)
            impl.set_wd(wd)
        r1, l1 = ch.compute2DL()
        r2, _, l2 = o.compute2DL()
        assert all(same(a, b) for a, b in zip(r1, r2)) and l1 == l2


def test_junction_boundaries_against_channet():
    """ChanNet's junction edges (ChanNet/compute2DL.c:154-161) take bzeta and
    bQn as given, with no sentinels: the oracle with bc_sentinels=0 against
    ChanNet's compiled compute2DL on meshes with prescribed edges whose
    values include negative surfaces and |qn| > 1000, on wet elements."""
    for seed in range(3):
        m = perturbed_mesh(N=8, seed=seed)
        st = random_state(m, seed=seed + 60, dz=0.2, q=0.3)
        rng = np.random.default_rng(seed)
        bz = rng.uniform(-0.5, 0.5, m.num_edges)
        bq = rng.uniform(-2000.0, 2000.0, m.num_edges)
        ch = oracle.RefChan(m)
        o = oracle.Oracle(m, math=oracle.MATH_REF, wetdry=1, bc_sentinels=0)
        wd = np.ones(m.num_el, np.int32)
        for impl in (ch, o):
            impl.set_state(*st)
            impl.set_wd(wd)
        ch.a["bz"][:] = bz
        ch.a["bq"][:] = bq
        o.set_boundary(bz, bq)
        assert np.any(np.array(m.bdry) > 0)
        r1, l1 = ch.compute2DL()
        r2, _, l2 = o.compute2DL()
        assert all(same(a, b) for a, b in zip(r1, r2)) and l1 == l2


@pytest.mark.parametrize("seed", range(3))
def test_unstructured_mesh_steps(seed):
    """Delaunay meshes (shuffled numbering, valences 1 to 10, every boundary
    code): compute2DL, the limiter and 12 wet/dry steps with boundary
    forcing, bit for bit against the compiled 2DCode and, for the wet/dry
    flux branches, ChanNet."""
    m = delaunay_mesh(seed=seed)
    st = random_state(m, seed=seed + 70, dz=0.1, q=0.1)
    ref, o = oracle.Ref2D(m), oracle.Oracle(m, math=oracle.MATH_REF)
    for impl in (ref, o):
        impl.set_state(*st)
    r1, f1, l1 = ref.compute2DL()
    r2, f2, l2 = o.compute2DL()
    assert all(same(a, b) for a, b in zip(f1 + r1, f2 + r2)) and l1 == l2
    ref.SlopeLimiter()
    o.SlopeLimiter()
    assert all(same(a, b) for a, b in zip(ref.get_state(), o.get_state()))
    for _ in range(12):
        ref.oneTimeStep2D(0.0, 0.05)
        o.oneTimeStep2D(0.0, 0.05)
    assert all(same(a, b) for a, b in zip(ref.get_state(), o.get_state()))
    mw = delaunay_mesh(seed=seed, codes=False)
    wd = np.random.default_rng(seed).integers(0, 2, mw.num_el).astype(np.int32)
    ch, ow = oracle.RefChan(mw), oracle.Oracle(mw, math=oracle.MATH_REF, wetdry=1)
    for impl in (ch, ow):
        impl.set_state(*random_state(mw, seed=seed + 80, dz=0.2, q=0.3))
        impl.set_wd(wd)
    r1, l1 = ch.compute2DL()
    r2, _, l2 = ow.compute2DL()
    assert all(same(a, b) for a, b in zip(r1, r2)) and l1 == l2


def test_topology_random_meshes(tmp_path):
    """The O(N) topology against the reference's extract_mesh_attributes on
    the same fort.14: grids, a Delaunay mesh, and a Delaunay mesh renumbered
    by locality_order and written back as a fort.14 (boundary sections
    included)."""
    from dgrain import mesh as M
    r = M.locality_order(delaunay_mesh(seed=6))[0]
    renumbered = M.from_elements(r.x, r.y, r.z, r.n, r.elto_vert, r.in1, r.in2, r.out)
    for N, m in ((4, perturbed_mesh(N=4, seed=4)), (9, perturbed_mesh(N=9, seed=9)), (0, delaunay_mesh(seed=5)),
                 (1, renumbered)):
        p = str(tmp_path / ("m%d.14" % N))
        M.write_fort14(p, m)
        t = oracle.Ref2D(m).topology_from_fort14(p)
        assert np.array_equal(t["edg_to_vert"], m.edg_to_vert)
        assert np.array_equal(t["edg_to_els"], m.edg_to_els)
        assert np.array_equal(t["bdry"], m.bdry)
        assert t["min_edge_length"] == m.min_edge_length


@pytest.mark.parametrize("seed", range(3))
def test_channel_coupling_against_reference(seed):
    """The junction side of ChanNet's internal_BC: coupled edges' bzeta/bQn
    and the ends' ghost area and discharge, bit for bit."""
    meshes, states, spec = coupling_network(seed)
    refs, ors = [oracle.RefChan(m) for m in meshes], [oracle.Oracle(m, math=oracle.MATH_REF) for m in meshes]
    for r, o, st, m in zip(refs, ors, states, meshes):
        r.set_state(*st)
        o.set_state(*st)
        o.set_boundary(np.zeros(m.num_edges), np.zeros(m.num_edges))   # RefChan starts at 0
    ga_r, gq_r = oracle.RefChan.internal_bc(refs, *spec)
    ga_o, gq_o = oracle.couple_ends(ors, *spec)
    assert same(ga_r, ga_o) and same(gq_r, gq_o)
    for r, o in zip(refs, ors):
        for a, b in zip(r.get_boundary(), o.get_boundary()):
            assert same(a, b)
    assert np.any(ors[0].get_boundary()[0] < 0)   # a below-datum surface: no sentinel applies here
