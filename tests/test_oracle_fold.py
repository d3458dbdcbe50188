"""The max-|lambda| fold over both-dry edges with NaN operands, checked by
brute force.

On a mesh where every element is flagged dry, every edge takes the
reference's both-dry branch (2DCode/compute2DL.c:203-238): two RoeFlux
evaluations with reflected states and localG = 0, folded into max_lambda
with the `max` macro of 2DCode/mathfunctions.h:8-10, ((a) > (b) ? (a) : (b)),
the first edge as max(in, ex) and every later one as max(max(m, in), ex).
The reference never checks these operands for NaN; a zero depth divides
0/0 and a NaN operand makes the fold restart at the next operand.

For still water (q = 0) the six |lambda| values RoeFlux folds
(numericalFlux2D.c:62-72,101-105) are closed forms of the depth alone:
|u n - sqrt(gH)| = sqrt(gH) twice per side, |u n| = 0, and sqrt(gH) again,
with u = 0/H = NaN at H = 0.  The test folds those operands in edge order
with the macro, literally, and compares with the oracle's compute2DL (which
folds per thread segment and joins the segments), on a grid, on a shuffled
mesh, and on that mesh renumbered with its edges sorted for locality.
"""
import math

import numpy as np
import pytest

import oracle
from dgrain import configs
from dgrain import mesh as M


def _max(a, b):
    # 2DCode/mathfunctions.h:8-10
    return a if a > b else b


def _roe_lam_still(H, g=9.81):
    """RoeFlux's returned max over its six |lambda| (numericalFlux2D.c:101-105)
    for equal still states of depth H on both sides."""
    u = 0.0 / H if H != 0 else math.nan      # Qx / H
    s = math.sqrt(g * H)
    l1 = abs(u - s)
    l2 = abs(u)
    l3 = abs(u + s)
    m = _max(l1, l1)
    m = _max(m, l2)
    m = _max(m, l2)
    m = _max(m, l3)
    return _max(m, l3)


def _brute_fold(mesh, h_el, zb):
    e2e = np.asarray(mesh.edg_to_els).reshape(-1, 2)
    m = None
    for i, (a, b) in enumerate(e2e):
        lin = _roe_lam_still((-zb + h_el[a]) + zb)
        lex = _roe_lam_still((-zb + h_el[b]) + zb)
        if i == 0:
            m = _max(lin, lex)
        else:
            m = _max(m, lin)
            m = _max(m, lex)
    return m


def _dry_flat(mesh, zb, rng, p_nan, last_nan=False):
    """Flat bed at depth zb; per element a film depth h (0 with probability
    p_nan, else 1e-5 .. 5e-4 m: below H0, so flagged dry) on all three
    midpoints; no flow."""
    ne = mesh.num_el
    h = rng.uniform(1e-5, 5e-4, ne)
    h[rng.random(ne) < p_nan] = 0.0
    if last_nan:
        a, b = np.asarray(mesh.edg_to_els).reshape(-1, 2)[-1]
        h[a] = h[b] = 0.0
    zeta = np.repeat(-zb + h, 3)
    return h, zeta


def _check(mesh, seed, p_nan, last_nan=False, zb=1.0):
    rng = np.random.default_rng(seed)
    h, zeta = _dry_flat(mesh, zb, rng, p_nan, last_nan)
    want = _brute_fold(mesh, h, zb)
    for threads in (1, 4):
        o = oracle.Oracle(mesh, math=oracle.MATH_DEV, wetdry=1, threads=threads)
        o.set_state(zeta, np.zeros_like(zeta), np.zeros_like(zeta))
        o.set_wd(np.zeros(mesh.num_el, np.int32))
        _, _, got = o.compute2DL()
        o.close()
        assert (math.isnan(want) and math.isnan(got)) or got == want, (threads, got, want)
    return want


def _flat(mesh, zb):
    return M.from_elements(mesh.x, mesh.y, np.full(mesh.num_nodes, zb), mesh.n, mesh.elto_vert.reshape(-1, 3))


@pytest.mark.parametrize("p_nan", [0.0, 0.05, 0.3])
def test_fold_grid(p_nan):
    wl = configs.dam_break(N=10, nsteps=0)
    want = _check(_flat(wl.mesh, 1.0), seed=1, p_nan=p_nan)
    assert p_nan > 0 or not math.isnan(want)


def test_fold_restarts_after_nan():
    """Cases where the fold's restart matters: the result is finite and
    below the largest operand (a NaN-ignoring max such as fmax would give
    the largest)."""
    wl = configs.dam_break(N=10, nsteps=0)
    mesh = _flat(wl.mesh, 1.0)
    restarted = 0
    for seed in range(12):
        want = _check(mesh, seed=100 + seed, p_nan=0.2)
        h, _ = _dry_flat(mesh, 1.0, np.random.default_rng(100 + seed), 0.2)
        top = max(_roe_lam_still((-1.0 + x) + 1.0) for x in h if x > 0)
        restarted += (not math.isnan(want)) and want < top
    assert restarted >= 3


def test_fold_last_operand_nan():
    """A NaN as the very last operand leaves max_lambda NaN."""
    wl = configs.dam_break(N=8, nsteps=0)
    assert math.isnan(_check(_flat(wl.mesh, 1.0), seed=2, p_nan=0.1, last_nan=True))


def test_fold_shuffled_and_locality_sorted():
    """On a shuffled mesh and on its renumbering with edges sorted for
    locality (dgrain.mesh.locality_order(edges="locality")), the oracle's
    dt operand is the left fold in each table's own edge order; the two
    differ, which is why that renumbering is bit for bit the original run
    only without both-dry edges."""
    from test_locality_order import _shuffled
    from helpers import delaunay_mesh
    sh = _shuffled(_flat(delaunay_mesh(seed=8, n_in=300), 1.0), seed=4)
    m, rn = M.locality_order(sh, edges="locality")
    rng = np.random.default_rng(7)
    h = rng.uniform(1e-5, 5e-4, sh.num_el)
    h[rng.random(sh.num_el) < 0.1] = 0.0
    folds = []
    for mesh, hh in ((sh, h), (m, h[rn.el_order])):
        want = _brute_fold(mesh, hh, 1.0)
        o = oracle.Oracle(mesh, math=oracle.MATH_DEV, wetdry=1, threads=3)
        zeta = np.repeat(-1.0 + hh, 3)
        o.set_state(zeta, np.zeros_like(zeta), np.zeros_like(zeta))
        o.set_wd(np.zeros(mesh.num_el, np.int32))
        got = o.compute2DL()[2]
        o.close()
        assert got == want
        folds.append(got)
    assert folds[0] != folds[1]
