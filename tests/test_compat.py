"""libdgrain_compat.so: the reference's own operator names over the
reference's globals, driven by a C program (tests/compat_driver.c) that
links it the way the reference's main.c and mesh/initialisation files would.
Results are compared bit-for-bit with the oracle (device-math mode)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from dgrain import configs
from helpers import perturbed_mesh, random_state, same

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "bin", "compat_driver")

pytestmark = pytest.mark.gpu


def _write_input(path, mesh, state, wd, wetdry, cfl):
    with open(path, "wb") as f:
        np.array([mesh.num_el, mesh.num_nodes, mesh.num_edges, wetdry], np.int32).tofile(f)
        np.array([cfl], np.float64).tofile(f)
        for a in (mesh.x, mesh.y, mesh.z, mesh.n):
            np.asarray(a, np.float64).tofile(f)
        for a in (mesh.elto_vert, mesh.edg_to_vert, mesh.edg_to_els, mesh.bdry):
            np.asarray(a, np.int32).tofile(f)
        np.array([mesh.min_edge_length], np.float64).tofile(f)
        for a in state:
            np.asarray(a, np.float64).tofile(f)
        np.asarray(wd, np.int32).tofile(f)


def _run(tmp_path, mesh, state, wd, op, *args, wetdry=0, cfl=0.1):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/bin/compat_driver is not built; run __graft_entry__.build()")
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(inp, mesh, state, wd, wetdry, cfl)
    r = subprocess.run([DRIVER, inp, out, op] + [str(a) for a in args], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ne3 = 3 * mesh.num_el
    raw = np.fromfile(out, np.uint8)
    d = raw[: 8 * (9 * ne3 + 1)].view(np.float64)
    wd_out = raw[8 * (9 * ne3 + 1):].view(np.int32)
    st = [d[i * ne3:(i + 1) * ne3] for i in range(9)]
    return st[:3], st[3:6], st[6:9], d[-1], wd_out, r.stdout


def test_compat_compute2DL_and_limiter(gpu, tmp_path):
    m = perturbed_mesh(N=8, seed=5)
    st = random_state(m, seed=6)
    wd = np.full(m.num_el, -1, np.int32)
    _, rhs, fh, ml, _, _ = _run(tmp_path, m, st, wd, "rhs")
    o = oracle.Oracle(m, math=oracle.MATH_DEV)
    o.set_state(*st)
    r2, f2, l2 = o.compute2DL()
    assert all(same(a, b) for a, b in zip(rhs, r2)) and all(same(a, b) for a, b in zip(fh, f2)) and ml == l2
    s1, _, _, _, _, _ = _run(tmp_path, m, st, wd, "limit")
    o.set_state(*st)
    o.SlopeLimiter()
    assert all(same(a, b) for a, b in zip(s1, o.get_state()))


def test_compat_one_time_step(gpu, tmp_path):
    m = perturbed_mesh(N=8, seed=7)
    st = random_state(m, seed=8, dz=0.05, q=0.05)
    wd = np.full(m.num_el, -1, np.int32)
    s1, _, _, ml, _, _ = _run(tmp_path, m, st, wd, "steps", 6, 0.02)
    o = oracle.Oracle(m, math=oracle.MATH_DEV)
    o.set_state(*st)
    for _ in range(6):
        o.oneTimeStep2D(0.0, 0.02)
    assert all(same(a, b) for a, b in zip(s1, o.get_state()))
    assert ml == o.max_lambda


def test_compat_time_evolution_writes_reference_files(gpu, tmp_path):
    wl = configs.dam_break(N=12, L=1200.0)
    wd = np.full(wl.num_el, -1, np.int32)
    T = 30.0
    s1, _, _, _, _, out = _run(tmp_path, wl.mesh, (wl.zeta, wl.qx, wl.qy), wd, "evolve", T, cfl=0.1)
    o = oracle.Oracle(wl.mesh, math=oracle.MATH_DEV, cfl=0.1, final_time=T)
    o.set_state(wl.zeta, wl.qx, wl.qy)
    while o.t < T:
        o.run(1)
    assert o.t == T
    assert all(same(a, b) for a, b in zip(s1, o.get_state()))
    assert out.count("dt = ") > 1
    for name in ("000_zeta.dat", "000_Qx.dat", "000_Qy.dat", "001_H.dat", "001_Qx.dat", "001_Qy.dat"):
        assert os.path.exists(str(tmp_path / name)), name
    vals = np.loadtxt(str(tmp_path / "001_H.dat")).ravel()
    assert np.allclose(vals, o.get_state()[0], rtol=0, atol=1e-15)

