"""dgr_locality_order and dgrain.mesh.locality_order: a renumbering of a mesh
for memory locality.  The renumbered mesh is the same mesh (same triangles,
boundary sections and codes), its numbering follows space (neighbours get
nearby numbers), and a run on it is the run on the original mesh, bit for
bit, in the oracle (the CPU restatement of the reference step)."""
import dataclasses

import numpy as np
import pytest

import oracle
from dgrain import configs
from dgrain import mesh as M
from dgrain import native
from helpers import delaunay_mesh, perturbed_mesh, same


def _shuffled(mesh, seed=0):
    """The mesh with its node and element numbers shuffled."""
    rng = np.random.default_rng(seed)
    pn = rng.permutation(mesh.num_nodes)
    inv = np.empty_like(pn)
    inv[pn] = np.arange(pn.size)
    ev = inv[mesh.elto_vert.reshape(-1, 3)[rng.permutation(mesh.num_el)]]
    lists = [inv[np.asarray(a)] if np.size(a) else None for a in (mesh.in1, mesh.in2, mesh.out)]
    return M.from_elements(mesh.x[pn], mesh.y[pn], mesh.z[pn], mesh.n[pn], ev, *lists)


def _triangles(mesh):
    ev = mesh.elto_vert.reshape(-1, 3)
    return sorted(tuple(sorted(zip(mesh.x[t], mesh.y[t]))) for t in ev)


def _neighbour_spread(mesh):
    """Mean |i - j| over the element pairs that share an edge."""
    e = np.asarray(mesh.edg_to_els).reshape(-1, 2)
    e = e[e[:, 0] != e[:, 1]]
    return float(np.abs(e[:, 0] - e[:, 1]).mean())


@pytest.mark.parametrize("make", [lambda: perturbed_mesh(N=24, seed=3), lambda: delaunay_mesh(seed=4, n_in=400)])
def test_same_mesh_better_locality(make):
    base = make()
    sh = _shuffled(base, seed=1)
    m, rn = M.locality_order(sh)
    assert np.array_equal(np.sort(rn.el_order), np.arange(sh.num_el))
    assert np.array_equal(np.sort(rn.node_order), np.arange(sh.num_nodes))
    assert (m.num_el, m.num_nodes, m.num_edges) == (sh.num_el, sh.num_nodes, sh.num_edges)
    assert _triangles(m) == _triangles(sh)
    # every edge the same segment between the same two triangles, same code
    ev_new, ev_old = m.elto_vert.reshape(-1, 3), sh.elto_vert.reshape(-1, 3)
    for k in (0, 1):
        a = np.asarray(m.edg_to_els)[k::2]
        b = np.asarray(sh.edg_to_els)[k::2]
        assert np.array_equal(rn.el_order[a], b)
        assert np.array_equal(m.x[ev_new[a]], sh.x[ev_old[b]])
    pa, pb = np.asarray(m.edg_to_vert).reshape(-1, 2), np.asarray(sh.edg_to_vert).reshape(-1, 2)
    assert np.array_equal(np.sort(m.x[pa] + 1j * m.y[pa], axis=1), np.sort(sh.x[pb] + 1j * sh.y[pb], axis=1))
    assert np.array_equal(m.bdry, sh.bdry)
    assert np.all(np.diff(np.asarray(m.edg_to_vert).reshape(-1, 2), axis=1) > 0)   # v0 < v1
    # the table is the one the reference builds for the renumbered elements,
    # up to edge order and orientation
    rebuilt = M.from_elements(m.x, m.y, m.z, m.n, m.elto_vert, m.in1, m.in2, m.out)
    key = lambda mm: sorted((min(p), max(p), c) for p, c in zip(np.asarray(mm.edg_to_vert).reshape(-1, 2).tolist(),
                                                                np.asarray(mm.bdry).tolist()))
    assert key(rebuilt) == key(m)
    assert np.array_equal(rn.nodes_to_new(sh.z), m.z) and np.array_equal(rn.nodes_to_old(m.n), sh.n)
    # nodes in order of first use by the new elements
    first = np.unique(m.elto_vert, return_index=True)[1]
    assert np.all(np.diff(first) >= 0)
    spread, spread_shuffled = _neighbour_spread(m), _neighbour_spread(sh)
    assert spread < 0.1 * spread_shuffled, (spread, spread_shuffled)


def test_run_on_renumbered_mesh_is_bitwise_the_same():
    """A shuffled Delaunay mesh and its renumbering, 30 wet/dry steps with
    tide and rain: states, wet/dry flags, limiter activations, t and dt equal
    after mapping back; and the same over 40 steps of a dam break onto a dry
    bed, where both-dry edges give NaN max-|lambda| operands."""
    sh = _shuffled(delaunay_mesh(seed=6, n_in=200), seed=2)
    m, rn = M.locality_order(sh)
    rng = np.random.default_rng(3)
    st = [0.05 * rng.standard_normal(3 * sh.num_el) for _ in range(3)]
    kw = dict(cfl=0.3, wetdry=1, tide_code=2, tide_datum=0.0, tide_amp=0.3, tide_period=600.0,
              rain=30.0 / 3.6e6)
    a = oracle.Oracle(sh, math=oracle.MATH_DEV, **kw)
    b = oracle.Oracle(m, math=oracle.MATH_DEV, **kw)
    a.set_state(*st)
    b.set_state(*[rn.midpoints_to_new(x) for x in st])
    ta = a.run(30)
    tb = b.run(30)
    assert ta == tb
    for x, y in zip(a.get_state(), b.get_state()):
        assert same(x, rn.midpoints_to_old(y))
    assert np.array_equal(a.get_wd(), rn.elements_to_old(b.get_wd()))
    for s in (0, 1):
        assert np.array_equal(a.get_modify(s), rn.elements_to_old(b.get_modify(s)))
    assert np.array_equal(rn.element_ids_to_new(rn.el_order), np.arange(sh.num_el))
    # dam break onto a dry flat bed (both-dry edges: NaN max-|lambda| operands)
    d = delaunay_mesh(seed=9, n_in=200, L=10.0, codes=False)
    d = dataclasses.replace(d, z=np.zeros(d.num_nodes))
    sh = _shuffled(d, seed=3)
    m, rn = M.locality_order(sh)
    z0 = M.nodal_to_midpoints(sh, np.where(sh.x <= 5.0, 0.005, configs.H0))
    st = (z0, np.zeros_like(z0), np.zeros_like(z0))
    a = oracle.Oracle(sh, math=oracle.MATH_DEV, cfl=0.3, wetdry=1)
    b = oracle.Oracle(m, math=oracle.MATH_DEV, cfl=0.3, wetdry=1)
    a.set_state(*st)
    b.set_state(*[rn.midpoints_to_new(x) for x in st])
    assert a.run(40) == b.run(40)
    for x, y in zip(a.get_state(), b.get_state()):
        assert same(x, rn.midpoints_to_old(y))
    wd = a.get_wd()
    assert np.array_equal(wd, rn.elements_to_old(b.get_wd())) and (wd == 0).any() and (wd == 1).any()


def test_errors():
    x = np.array([0.0, 1.0, 0.0])
    y = np.array([0.0, 0.0, 1.0])
    with pytest.raises(native.DgrError):
        native.locality_order(np.array([0, 1, 3], np.int32), x, y)
    with pytest.raises(native.DgrError):
        native.locality_order(np.array([0, 1, 2], np.int32), np.array([0.0, np.nan, 0.0]), y)
    eo, no = native.locality_order(np.zeros(0, np.int32), x, y)
    assert eo.size == 0 and np.array_equal(no, [0, 1, 2])


def test_locality_edge_order():
    """edges="locality": edges sorted by their first new element, each with
    its orientation and code; a wet run with tide and rain (no NaN operand
    in the max-|lambda| fold) is the original run, bit for bit."""
    sh = _shuffled(delaunay_mesh(seed=8, n_in=300), seed=4)
    m, rn = M.locality_order(sh, edges="locality")
    e_new = np.asarray(m.edg_to_els).reshape(-1, 2)
    e_old = np.asarray(sh.edg_to_els).reshape(-1, 2)
    assert np.all(np.diff(e_new.min(axis=1)) >= 0)
    assert np.array_equal(rn.el_order[e_new], e_old[rn.edge_order])
    assert np.array_equal(m.bdry, rn.edges_to_new(sh.bdry))
    assert np.array_equal(rn.edges_to_old(m.bdry), sh.bdry)
    rng = np.random.default_rng(5)
    st = [0.05 * rng.standard_normal(3 * sh.num_el) for _ in range(3)]
    kw = dict(cfl=0.3, wetdry=1, tide_code=2, tide_datum=0.0, tide_amp=0.3, tide_period=600.0, rain=30.0 / 3.6e6)
    a = oracle.Oracle(sh, math=oracle.MATH_DEV, **kw)
    b = oracle.Oracle(m, math=oracle.MATH_DEV, **kw)
    a.set_state(*st)
    b.set_state(*[rn.midpoints_to_new(x) for x in st])
    assert a.run(30) == b.run(30)
    for x, y in zip(a.get_state(), b.get_state()):
        assert same(x, rn.midpoints_to_old(y))
    with pytest.raises(ValueError):
        M.locality_order(sh, edges="sorted")
