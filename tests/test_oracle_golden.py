"""The oracle (reference-math mode) against the golden vectors that
tests/golden/make_golden.py wrote from the compiled reference.  Bit-for-bit.
These run anywhere: the vectors are data, the reference is not needed."""
import os

import numpy as np
import pytest

import oracle
from dgrain import configs, mesh as M
from helpers import same

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name + ".npz"))


def mesh_of(d):
    return M.Mesh(d["x"], d["y"], d["z"], d["n"], d["elto_vert"], d["edg_to_vert"], d["edg_to_els"], d["bdry"],
                  float(d["min_edge_length"]))


def test_roe_flux_vectors():
    d = load("roe_flux")
    for row, F, lam in zip(d["inputs"], d["F"], d["lam"]):
        f, l = oracle.roe_flux(*row)
        assert same(f, F) and l == lam, row


def test_one_call_operators():
    d = load("operators_6x6")
    m = mesh_of(d)
    o = oracle.Oracle(m, math=oracle.MATH_REF)
    o.set_state(d["zeta"], d["qx"], d["qy"])
    r, f, ml = o.compute2DL()
    assert same(np.stack(f), d["fhat"])
    assert same(np.stack(r), d["rhs"])
    assert ml == float(d["max_lambda"])
    for el in range(m.num_el):
        SI, VI = o.inner_products(el)
        assert same(SI, d["SI"][el]) and same(VI, d["VI"][el]), el
    o.SlopeLimiter()
    assert same(np.stack(o.get_state()), d["limited"])
    o.set_state(d["wd_state"], d["qx"], d["qy"])
    o.set_wd(d["wd_in"])
    o.wetDryStatus2D()
    assert np.array_equal(o.get_wd(), d["wd_out"])
    o.PDop2D()
    assert same(np.stack(o.get_state()), d["pdop"])


@pytest.mark.parametrize("name", ["traj_dam_break_8x8", "traj_prescribed_6x6", "traj_lake_at_rest_8x8"])
def test_2dcode_trajectories(name):
    d = load(name)
    m = mesh_of(d)
    o = oracle.Oracle(m, math=oracle.MATH_REF, wetdry=0, cfl=0.1)
    o.set_state(*d["init"])
    t, dt = 0.0, 1e-7
    for k, ref_state in enumerate(d["states"]):
        assert dt == d["dts"][k]
        o.run(1)
        assert same(np.stack(o.get_state()), ref_state), "%s step %d" % (name, k + 1)
        dt = o.dt
    assert dt == d["dts"][-1]


@pytest.mark.parametrize("name", ["traj_wetdry_beach_10x10", "traj_dry_bed_12x12"])
def test_wetdry_trajectories(name):
    d = load(name)
    m = mesh_of(d)
    o = oracle.Oracle(m, math=oracle.MATH_REF, wetdry=1, cfl=0.1)
    o.set_state(*d["init"])
    for k, (ref_state, ref_wd) in enumerate(zip(d["states"], d["wds"])):
        assert o.dt == d["dts"][k]
        o.run(1)
        assert same(np.stack(o.get_state()), ref_state), "%s step %d" % (name, k + 1)
        assert np.array_equal(o.get_wd(), ref_wd), "%s step %d" % (name, k + 1)


def test_topology_against_reference_extract_mesh_attributes():
    d = load("topology_5x5")
    m = M.read_fort14(os.path.join(G, "topology_5x5.14"))     # product O(N) builder
    assert m.num_edges == int(d["num_edges"])
    assert np.array_equal(m.edg_to_vert, d["edg_to_vert"])
    assert np.array_equal(m.edg_to_els, d["edg_to_els"])
    assert np.array_equal(m.bdry, d["bdry"])
    assert m.min_edge_length == float(d["min_edge_length"])
    e2v, e2e, bd, mel = oracle.build_topology(m.elto_vert, m.x, m.y, m.in1, m.in2, m.out)
    assert np.array_equal(e2v, d["edg_to_vert"]) and np.array_equal(e2e, d["edg_to_els"])
    assert np.array_equal(bd, d["bdry"]) and mel == float(d["min_edge_length"])


def test_still_water_coastal_fails_where_the_reference_fails():
    """The reference's wet/dry scheme stops on a NaN flux on the still-water
    coastal start; the oracle stops in the same step at the same edge."""
    d = load("still_coastal_failure")
    msg = str(d["message"])
    assert "not a number" in msg
    edge = int(msg.split("edge ")[1].split(",")[0])
    wl = configs.coastal(N=16, start="still")
    wl.mesh.bdry[:] = 0
    o = oracle.Oracle(wl.mesh, math=oracle.MATH_REF, wetdry=1, cfl=0.3)
    o.set_state(wl.zeta, wl.qx, wl.qy)
    for step in range(1, 50):
        try:
            o.run(1)
        except RuntimeError as ex:
            assert step == int(d["step"])
            assert str(ex).endswith("error 3 at %d" % edge)
            return
    pytest.fail("the oracle survived a start the reference cannot run")
