"""Shared test inputs: small perturbed meshes with every boundary code."""
import numpy as np

from dgrain import mesh as M


def perturbed_mesh(N=6, L=100.0, seed=1, depth=2.0, codes=True):
    """Grid with jittered interior nodes, random depth and n, and boundary
    sections of codes 1 (left, lower half), 3 (bottom) and 2 (right)."""
    rng = np.random.default_rng(seed)
    x, y = M.grid_nodes(N, L)
    inner = (x > 0) & (x < L) & (y > 0) & (y < L)
    h = L / N
    x = x + rng.uniform(-0.2 * h, 0.2 * h, x.size) * inner
    y = y + rng.uniform(-0.2 * h, 0.2 * h, y.size) * inner
    z = depth + rng.uniform(-0.25 * depth, 0.25 * depth, x.size)
    n = rng.uniform(0.01, 0.05, x.size)
    kw = {}
    if codes:
        kw = dict(in1=M.side_pairs(N, "left")[: 2 * (N // 2)], in2=M.side_pairs(N, "bottom"),
                  out=M.side_pairs(N, "right"))
    return M.from_elements(x, y, z, n, M.grid_elements(N), **kw)


def random_state(mesh, seed=2, dz=0.2, q=0.5):
    rng = np.random.default_rng(seed)
    ne3 = 3 * mesh.num_el
    return (rng.uniform(-dz, dz, ne3), rng.uniform(-q, q, ne3), rng.uniform(-q, q, ne3))


def beach_mesh(N=10, L=100.0, seed=3):
    """Sloping beach: depth +1 at x=0 to -0.5 at x=L (dry upper part)."""
    rng = np.random.default_rng(seed)
    x, y = M.grid_nodes(N, L)
    z = 1.0 - 1.5 * x / L + 0.05 * rng.uniform(-1, 1, x.size)
    n = rng.uniform(0.01, 0.05, x.size)
    return M.from_elements(x, y, z, n, M.grid_elements(N))


def beach_state(mesh, seed=4, H0=1e-3):
    rng = np.random.default_rng(seed)
    zn = np.maximum(0.2, H0 - mesh.z)
    zeta = M.nodal_to_midpoints(mesh, zn)
    Hm = zeta + M.nodal_to_midpoints(mesh, mesh.z)
    ne3 = 3 * mesh.num_el
    qx = rng.uniform(-0.02, 0.02, ne3) * (Hm > 0.1)
    qy = rng.uniform(-0.02, 0.02, ne3) * (Hm > 0.1)
    return zeta, qx, qy


def same(a, b):
    """Bitwise equality up to the sign of zero (NaN == NaN)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return d / n if n > 0 else d


def coupling_network(seed):
    """Two junction meshes and four channel ends: two inflow ends (one edge
    each, and three edges), two outflow ends (two edges on the first
    junction, one on the second), over boundary and interior edges, with
    random junction states (surface below and above the datum) and end
    values.  Returns (meshes, states, end spec)."""
    rng = np.random.default_rng(seed)
    meshes = [perturbed_mesh(N=5, seed=seed), perturbed_mesh(N=6, seed=seed + 1)]
    states = []
    for i, m in enumerate(meshes):
        z, qx, qy = random_state(m, seed=300 + 10 * seed + i, dz=0.3, q=0.2)
        states.append((z - 0.2, qx, qy))
    bd = [np.nonzero(np.array(m.edg_to_els[0::2]) == np.array(m.edg_to_els[1::2]))[0] for m in meshes]
    edges = [(0, int(bd[0][0]))], [(1, int(bd[1][3])), (1, int(bd[1][4])), (0, 7)], \
        [(0, int(bd[0][5])), (0, int(bd[0][6]))], [(1, int(bd[1][9]))]
    end_off = np.cumsum([0] + [len(e) for e in edges])
    edge_junc = np.array([j for e in edges for j, _ in e], np.int32)
    edge_local = np.array([k for e in edges for _, k in e], np.int32)
    end_dir = np.array([1, 1, -1, -1], np.int32)
    end_A = rng.uniform(0.5, 3.0, 4)
    end_Q = rng.uniform(-1.0, 1.0, 4)
    return meshes, states, (end_off, edge_junc, edge_local, end_dir, end_A, end_Q)


def delaunay_mesh(n_side=8, n_in=80, L=100.0, seed=0, depth=2.0, codes=True):
    """Unstructured mesh: the Delaunay triangulation of n_side + 1 evenly
    spaced nodes per side of the square and n_in random interior nodes (kept
    at least 0.4 of the mean spacing apart), with node and element numbers
    shuffled, every element counter-clockwise, node valences of 3 to ~10,
    random depth and n, and boundary sections of codes 1 (left, lower half),
    3 (bottom) and 2 (right)."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    t = np.arange(n_side) * (L / n_side)
    bx = np.concatenate([t, np.full(n_side, L), L - t, np.zeros(n_side)])
    by = np.concatenate([np.zeros(n_side), t, np.full(n_side, L), L - t])
    hmin = 0.4 * L / np.sqrt(n_in + 4 * n_side)
    px, py = list(bx), list(by)
    while len(px) < 4 * n_side + n_in:
        cx, cy = rng.uniform(0.05 * L, 0.95 * L, 2)
        if np.min(np.hypot(np.asarray(px) - cx, np.asarray(py) - cy)) >= hmin:
            px.append(cx)
            py.append(cy)
    x, y = np.asarray(px), np.asarray(py)
    tri = Delaunay(np.c_[x, y]).simplices.astype(np.int64)
    assert np.unique(tri).size == x.size, "a node was left out of the triangulation"
    area = (x[tri[:, 1]] - x[tri[:, 0]]) * (y[tri[:, 2]] - y[tri[:, 0]]) - \
        (x[tri[:, 2]] - x[tri[:, 0]]) * (y[tri[:, 1]] - y[tri[:, 0]])
    assert np.all(np.abs(area) > 1e-6 * hmin * hmin)
    tri[area < 0] = tri[area < 0][:, [0, 2, 1]]
    # shuffle node and element numbering
    perm = rng.permutation(x.size)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(x.size)
    x, y = x[perm], y[perm]
    tri = inv[tri][rng.permutation(tri.shape[0])]
    z = depth + rng.uniform(-0.25 * depth, 0.25 * depth, x.size)
    n = rng.uniform(0.01, 0.05, x.size)
    kw = {}
    if codes:
        def side(k0, k1):   # boundary nodes k0..k1 (in side order) as consecutive pairs
            ids = inv[np.arange(k0, k1 + 1) % (4 * n_side)]
            return np.ascontiguousarray(np.stack([ids[:-1], ids[1:]], 1), np.int32).ravel()
        kw = dict(in1=side(3 * n_side + n_side // 2, 4 * n_side), in2=side(0, n_side), out=side(n_side, 2 * n_side))
    return M.from_elements(x, y, z, n, tri.astype(np.int32).ravel(), **kw)


def smooth_field(x, y, rng, modes=4, kmax=3):
    """A smooth random field over points (x, y): a few cosine modes of low
    wave number over the bounding box."""
    x = (x - x.min()) / (np.ptp(x) + 1e-300)
    y = (y - y.min()) / (np.ptp(y) + 1e-300)
    d = np.zeros(np.size(x))
    for _ in range(modes):
        kx, ky = rng.integers(0, kmax, 2)
        d += rng.standard_normal() * np.cos(np.pi * kx * x + rng.uniform(0, 6.3)) * \
            np.cos(np.pi * ky * y + rng.uniform(0, 6.3))
    return d


def smooth_start(wl, rng, amp=0.05, qamp=0.05):
    """The workload started from its own state plus smooth random elevation
    and discharge fields (amp, qamp).  Unlike a per-midpoint random state,
    a smooth one leaves the limiter inactive almost everywhere, and unlike
    still water it has no exact ties in the limiter bounds or the max-|lambda|
    fold, so a loss of the run is differentiable at the start with few kinks
    nearby."""
    import dataclasses
    m = wl.mesh
    zn, un, vn = (a * smooth_field(m.x, m.y, rng) for a in (amp, qamp, qamp))
    return dataclasses.replace(wl, zeta=wl.zeta + M.nodal_to_midpoints(m, zn), qx=M.nodal_to_midpoints(m, un),
                               qy=M.nodal_to_midpoints(m, vn))


def converged_difference(loss_of, eps_list, tol, avg=1):
    """Central differences (L(+e) - L(-e)) / 2e for e in eps_list (largest
    first), each the mean over the avg step sizes e * (1 + 0.1 k), k < avg
    (the rounding noise of separate runs averages out).  Returns
    (estimate, e) at the largest e whose estimate agrees with the next
    smaller e's to within tol (relative): a kink of the piecewise-smooth loss
    inside [-e, e] would show as a disagreement.  Returns (None, None) if no
    successive pair agrees."""
    def fd(e):
        return sum((loss_of(x) - loss_of(-x)) / (2 * x) for x in (e * (1 + 0.1 * k) for k in range(avg))) / avg
    fds = [fd(e) for e in eps_list]
    for i in range(len(fds) - 1):
        a, b = fds[i], fds[i + 1]
        if abs(a - b) <= tol * max(abs(b), 1e-300):
            return b, eps_list[i + 1]
    return None, None
