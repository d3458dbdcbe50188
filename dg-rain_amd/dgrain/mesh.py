"""Meshes in the reference's fort.14 variant and the synthetic workloads.

Arrays follow the reference globals (2DCode/MeshAttributes.h:5-42):
x, y, z (depth, positive down), n (Manning, per node), EltoVert, EdgtoVert,
EdgtoEls, BdryPrescribed.  Topology comes from libdgrain's O(N) restatement of
extract_mesh_attributes() (2DCode/extract_mesh_attributes.c:143-361).
"""
from dataclasses import dataclass, field

import numpy as np

from . import native


@dataclass
class Mesh:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    n: np.ndarray
    elto_vert: np.ndarray
    edg_to_vert: np.ndarray
    edg_to_els: np.ndarray
    bdry: np.ndarray
    min_edge_length: float
    in1: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    in2: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    out: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))

    @property
    def num_el(self):
        return self.elto_vert.size // 3

    @property
    def num_nodes(self):
        return self.x.size

    @property
    def num_edges(self):
        return self.bdry.size

    def centroid_x(self):
        ev = self.elto_vert.reshape(-1, 3)
        return self.x[ev].mean(axis=1)


def from_elements(x, y, z, n, elto_vert, in1=None, in2=None, out=None):
    """Mesh from node arrays and EltoVert; topology by dgr_build_topology."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    ev = np.ascontiguousarray(elto_vert, np.int32).ravel()
    lists = [np.ascontiguousarray(a if a is not None else np.zeros(0), np.int32).ravel() for a in (in1, in2, out)]
    e2v, e2e, bdry, mel = native.build_topology(ev, x, y, *lists)
    return Mesh(x, y, np.ascontiguousarray(z, np.float64), np.ascontiguousarray(n, np.float64), ev,
                e2v, e2e, bdry, mel, *lists)


def grid_nodes(N, L, Ly=None, x0=0.0, y0=0.0):
    """(N+1)^2 row-major nodes a = j(N+1)+i over [x0,x0+L] x [y0,y0+Ly]."""
    Ly = L if Ly is None else Ly
    i = np.tile(np.arange(N + 1), N + 1)
    j = np.repeat(np.arange(N + 1), N + 1)
    return x0 + L * i / N, y0 + Ly * j / N


def grid_elements(N):
    """Cell (i,j) split into (a,b,c) and (a,c,d): a=j(N+1)+i, b=a+1, c=a+N+2, d=a+N+1."""
    j, i = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    a = (j * (N + 1) + i).ravel()
    b, c, d = a + 1, a + N + 2, a + N + 1
    ev = np.empty((2 * N * N, 3), np.int32)
    ev[0::2] = np.stack([a, b, c], 1)
    ev[1::2] = np.stack([a, c, d], 1)
    return ev.ravel()


def side_pairs(N, side):
    """Node pairs along one side of the N x N grid ('left','right','bottom','top')."""
    k = np.arange(N)
    if side == "left":
        p = np.stack([k * (N + 1), (k + 1) * (N + 1)], 1)
    elif side == "right":
        p = np.stack([k * (N + 1) + N, (k + 1) * (N + 1) + N], 1)
    elif side == "bottom":
        p = np.stack([k, k + 1], 1)
    elif side == "top":
        p = np.stack([N * (N + 1) + k, N * (N + 1) + k + 1], 1)
    else:
        raise ValueError(side)
    return np.ascontiguousarray(p, np.int32).ravel()


def grid_mesh(N, L, z=0.0, n=0.02, Ly=None, x0=0.0, y0=0.0, in1=None, in2=None, out=None):
    x, y = grid_nodes(N, L, Ly, x0, y0)
    zz = np.broadcast_to(np.asarray(z, np.float64), x.shape).copy() if np.ndim(z) == 0 else np.asarray(z, np.float64)
    nn = np.broadcast_to(np.asarray(n, np.float64), x.shape).copy() if np.ndim(n) == 0 else np.asarray(n, np.float64)
    return from_elements(x, y, zz, nn, grid_elements(N), in1, in2, out)


def nodal_to_midpoints(mesh, nodal):
    """Midpoint values 0.5(v_k + v_{k+1}) at el*3+k (initialize.c:117-123)."""
    ev = mesh.elto_vert.reshape(-1, 3)
    v = np.asarray(nodal, np.float64)[ev]
    out = np.empty_like(v)
    out[:, 0] = 0.5 * (v[:, 0] + v[:, 1])
    out[:, 1] = 0.5 * (v[:, 1] + v[:, 2])
    out[:, 2] = 0.5 * (v[:, 2] + v[:, 0])
    return out.ravel()


@dataclass
class Renumbering:
    """A mesh renumbered by locality_order: new element i is old element
    el_order[i], new node j is old node node_order[j], new edge e is old
    edge edge_order[e]."""
    el_order: np.ndarray
    node_order: np.ndarray
    edge_order: np.ndarray

    def edges_to_new(self, a):
        """Per-edge array (boundary values) of the old numbering in the new one."""
        return np.ascontiguousarray(np.asarray(a)[self.edge_order])

    def edges_to_old(self, a):
        out = np.empty_like(np.asarray(a))
        out[self.edge_order] = a
        return out

    def midpoints_to_new(self, a):
        """Per-midpoint array (el*3+k) of the old numbering in the new one."""
        return np.ascontiguousarray(np.asarray(a).reshape(-1, 3)[self.el_order].ravel())

    def midpoints_to_old(self, a):
        out = np.empty_like(np.asarray(a).reshape(-1, 3))
        out[self.el_order] = np.asarray(a).reshape(-1, 3)
        return out.ravel()

    def elements_to_new(self, a):
        return np.ascontiguousarray(np.asarray(a)[self.el_order])

    def elements_to_old(self, a):
        out = np.empty_like(np.asarray(a))
        out[self.el_order] = a
        return out

    def element_ids_to_new(self, ids):
        """New numbers of old elements (gauge lists and the like)."""
        inv = np.empty_like(self.el_order)
        inv[self.el_order] = np.arange(self.el_order.size, dtype=self.el_order.dtype)
        return inv[np.asarray(ids)]

    def nodes_to_new(self, a):
        return np.ascontiguousarray(np.asarray(a)[self.node_order])

    def nodes_to_old(self, a):
        out = np.empty_like(np.asarray(a))
        out[self.node_order] = a
        return out


def locality_order(mesh, edges="keep"):
    """The mesh renumbered for memory locality (dgr_locality_order: elements
    along a Hilbert curve through their centroids, nodes in order of first
    use); returns (mesh, Renumbering).  Each edge keeps its orientation
    (which element is el_in), with element and node numbers mapped: the
    reference evaluates each Roe flux in its edge's orientation, so
    rebuilding the table from the new numbering would change results at the
    rounding level.

    edges="keep": the edge table also keeps its order.  A run on the
    renumbered mesh is then the original run, renumbered, bit for bit.
    edges="locality": edges ordered by the first new element that has them
    (the edge-centric kernels then gather with locality too), orientation
    kept.  Without both-dry edges the run is the original run, bit for bit.
    The reference restarts its max-|lambda| fold at NaN operands (both-dry
    edges) in edge order, so a step that has them takes the dt of this
    table's own fold order.  That is neither the original run's dt nor the
    dt of the reference's run of a renumbered fort.14: extract_mesh_attributes
    would order the edges by their lowest vertex and make the lowest-numbered
    element el_in.  The Renumbering holds the edge order too (new edge e is
    old edge edge_order[e]; the identity for "keep")."""
    if edges not in ("keep", "locality"):
        raise ValueError(edges)
    el_order, node_order = native.locality_order(mesh.elto_vert, mesh.x, mesh.y)
    inv_n = np.empty_like(node_order)
    inv_n[node_order] = np.arange(node_order.size, dtype=node_order.dtype)
    inv_e = np.empty_like(el_order)
    inv_e[el_order] = np.arange(el_order.size, dtype=el_order.dtype)

    def nodes(a):
        return np.ascontiguousarray(inv_n[np.asarray(a, np.int64)], np.int32) if np.size(a) else \
            np.zeros(0, np.int32)

    e2e = inv_e[np.asarray(mesh.edg_to_els, np.int64).reshape(-1, 2)]
    edge_order = np.arange(e2e.shape[0], dtype=np.int32)
    if edges == "locality":
        edge_order = np.argsort(e2e.min(axis=1), kind="stable").astype(np.int32)
    # each edge's vertex pair in ascending new numbers, as the reference
    # stores it (its length and z_edge are symmetric in the two vertices)
    e2v = np.sort(inv_n[np.asarray(mesh.edg_to_vert, np.int64).reshape(-1, 2)[edge_order]], axis=1)
    m = Mesh(mesh.x[node_order], mesh.y[node_order], mesh.z[node_order], mesh.n[node_order],
             nodes(mesh.elto_vert.reshape(-1, 3)[el_order].ravel()), np.ascontiguousarray(e2v.ravel(), np.int32),
             np.ascontiguousarray(e2e[edge_order].ravel(), np.int32),
             np.ascontiguousarray(np.asarray(mesh.bdry)[edge_order], np.int32), mesh.min_edge_length,
             nodes(mesh.in1), nodes(mesh.in2), nodes(mesh.out))
    return m, Renumbering(el_order, node_order, edge_order)


# ---------------------------------------------------------------------------
# fort.14 variant (extract_mesh_attributes.c:39-141)
# ---------------------------------------------------------------------------
def write_fort14(path, mesh, title="dgrain mesh"):
    with open(path, "w") as f:
        f.write(title + "\n")
        f.write("%d %d\n" % (mesh.num_el, mesh.num_nodes))
        for i in range(mesh.num_nodes):
            f.write("%d %.17g %.17g %.17g %.17g\n" % (i + 1, mesh.x[i], mesh.y[i], mesh.z[i], mesh.n[i]))
        ev = mesh.elto_vert.reshape(-1, 3)
        for e in range(mesh.num_el):
            f.write("%d 3 %d %d %d\n" % (e + 1, ev[e, 0] + 1, ev[e, 1] + 1, ev[e, 2] + 1))
        f.write("boundaries\n")
        for lst in (mesh.in1, mesh.in2, mesh.out):
            p = np.asarray(lst).reshape(-1, 2)
            f.write("%d\n" % len(p))
            for a, b in p:
                f.write("%d %d\n" % (a + 1, b + 1))


def read_fort14(path):
    """Validating reader of the fort.14 variant; raises ValueError on short sections."""
    with open(path) as f:
        lines = f.read().splitlines()
    pos = 1
    ne, nn = (int(t) for t in lines[pos].split()[:2])
    pos += 1
    nodes = np.array([[float(t) for t in lines[pos + i].split()[1:5]] for i in range(nn)])
    pos += nn
    ev = np.array([[int(t) for t in lines[pos + i].split()[2:5]] for i in range(ne)], np.int32) - 1
    pos += ne + 1
    lists = []
    for _ in range(3):
        if pos >= len(lines):
            raise ValueError("fort.14: missing boundary section")
        cnt = int(lines[pos].split()[0])
        pos += 1
        if pos + cnt > len(lines):
            raise ValueError("fort.14: short boundary section")
        lists.append(np.array([[int(t) for t in lines[pos + i].split()[:2]] for i in range(cnt)], np.int32).reshape(-1) - 1)
        pos += cnt
    return from_elements(nodes[:, 0], nodes[:, 1], nodes[:, 2], nodes[:, 3], ev.ravel(), *lists)
