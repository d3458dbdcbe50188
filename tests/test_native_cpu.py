"""CPU-side checks of the product library (no kernel launches)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from dgrain import mesh as M, native
from helpers import perturbed_mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = native.load()
    names = native.exported_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_params_default_are_reference_constants():
    p = native.default_params()
    assert (p.g, p.h0, p.velzero, p.cfl, p.dt0) == (9.81, 1e-3, 1e-2, 0.1, 1e-7)   # initialize.c:22-27, time_evolution.c:255,399
    assert p.wetdry == 0 and p.tide_code == 0 and p.rain == 0.0


@pytest.mark.parametrize("N", [3, 7, 20])
def test_topology_matches_oracle(N):
    m = perturbed_mesh(N=N, seed=N)
    e2v, e2e, bd, mel = oracle.build_topology(m.elto_vert, m.x, m.y, m.in1, m.in2, m.out)
    assert np.array_equal(e2v, m.edg_to_vert)
    assert np.array_equal(e2e, m.edg_to_els)
    assert np.array_equal(bd, m.bdry)
    assert mel == m.min_edge_length
    assert m.num_edges == m.num_nodes + m.num_el - 1


def test_topology_rejects_non_simply_connected():
    # two disjoint triangles: Euler's count NumEdges = NumNodes + NumEl - 1 fails
    ev = np.array([0, 1, 2, 3, 4, 5], np.int32)
    x = np.array([0, 1, 0, 5, 6, 5.0])
    y = np.array([0, 0, 1, 0, 0, 1.0])
    with pytest.raises(native.DgrError):
        native.build_topology(ev, x, y)


def test_fort14_roundtrip(tmp_path):
    m = perturbed_mesh(N=5, seed=2)
    p = os.path.join(tmp_path, "m.14")
    M.write_fort14(p, m)
    m2 = M.read_fort14(p)
    for a in ("x", "y", "z", "n", "elto_vert", "edg_to_vert", "edg_to_els", "bdry"):
        assert np.array_equal(getattr(m, a), getattr(m2, a)), a


def test_fort14_short_section_is_an_error(tmp_path):
    m = perturbed_mesh(N=3, seed=2)
    p = os.path.join(tmp_path, "m.14")
    M.write_fort14(p, m)
    lines = open(p).read().splitlines()
    open(p, "w").write("\n".join(lines[:-3]) + "\n")
    with pytest.raises(ValueError):
        M.read_fort14(p)


def test_create_without_device_reports_nodev():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    m = perturbed_mesh(N=3)
    with pytest.raises(native.DgrError) as ei:
        native.Solver(m)
    assert ei.value.code == 6


def test_create_rejects_bad_mesh():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("mesh validation runs inside dgr_create after the device check")
    m = perturbed_mesh(N=3)
    bad = M.Mesh(m.x, m.y, m.z, m.n, m.elto_vert, m.edg_to_vert.copy(), m.edg_to_els.copy(), m.bdry,
                 m.min_edge_length)
    bad.edg_to_els[0] = bad.edg_to_els[1] = 5   # edge 0 does not belong to element 5
    with pytest.raises(native.DgrError) as ei:
        native.Solver(bad)
    assert ei.value.code == 2


def test_num_variants_matches_header():
    hdr = open(os.path.join(ROOT, "include", "dgrain.h")).read()
    m = re.search(r"#define DGR_NUM_VARIANTS (\d+)", hdr)
    assert m and int(m.group(1)) == native.NUM_VARIANTS
