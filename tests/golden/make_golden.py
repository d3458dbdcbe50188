"""Write the golden vectors of tests/golden/ from the compiled reference.

Runs only where the reference sources are mounted: it builds oracle/_ref from
them (oracle/Makefile) and drives the reference's own operators:
  * 2DCode (libref2dcode.so): RoeFlux, Compute2DInnerProducts, compute2DL,
    SlopeLimiter, wetDryStatus2D, PDop2D, extract_mesh_attributes, and the
    step loop of time_evolution.c:311-402 composed over them;
  * ChanNet built -DWDON (librefchan2d.so): the wet/dry step of
    ChanNet/time_evolution.c:218-427.
Outputs are inputs plus expected outputs only (.npz and one fort.14 mesh).

    python tests/golden/make_golden.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "dg-rain_amd"), os.path.join(ROOT, "tests")]

import oracle  # noqa: E402
from dgrain import configs, mesh as M  # noqa: E402
from helpers import beach_mesh, beach_state, perturbed_mesh, random_state  # noqa: E402


def mesh_dict(m):
    return dict(x=m.x, y=m.y, z=m.z, n=m.n, elto_vert=m.elto_vert, edg_to_vert=m.edg_to_vert,
                edg_to_els=m.edg_to_els, bdry=m.bdry, min_edge_length=np.array(m.min_edge_length))


def roe_cases(rng, n=400):
    rows = []
    for i in range(n):
        zb = rng.uniform(0.5, 5.0)
        zin, zex = rng.uniform(-0.4, 0.6, 2)
        qxin, qxex, qyin, qyex = rng.uniform(-2, 2, 4)
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(th), np.sin(th)
        lg = [9.81, 0.0][i % 2]
        if i % 7 == 0:   # consistency cases: identical states
            zex, qxex, qyex = zin, qxin, qyin
        rows.append([zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, lg])
    return np.array(rows)


def ref_dt_run(ref, nsteps, cfl=0.1, dt0=1e-7):
    """time_evolution.c:311-402 over the reference operators: states after
    every step and the dt sequence."""
    dt = dt0
    states, dts = [], []
    for _ in range(nsteps):
        ref.oneTimeStep2D(0.0, dt)
        states.append(np.stack(ref.get_state()))
        dts.append(dt)
        dt = cfl * ref.mesh.min_edge_length / ref.max_lambda
    return np.stack(states), np.array(dts + [dt])


def chan_run(ch, nsteps, cfl=0.1, dt0=1e-7):
    dt = dt0
    states, wds, dts = [], [], []
    for _ in range(nsteps):
        ml = ch.oneTimeStep2D(0.0, dt)
        states.append(np.stack(ch.get_state()))
        wds.append(ch.get_wd())
        dts.append(dt)
        dt = cfl * ch.mesh.min_edge_length / ml
    return np.stack(states), np.stack(wds), np.array(dts + [dt])


def still_coastal_failure(N=16):
    """Run ChanNet -DWDON on the 'still water' coastal start (walls only) in a
    child process; the reference prints its NaN message and exits."""
    code = r'''
import sys, numpy as np
sys.path[:0] = %r
import oracle
from dgrain import configs
wl = configs.coastal(N=%d, start="still")
wl.mesh.bdry[:] = 0
ch = oracle.RefChan(wl.mesh)
ch.set_state(wl.zeta, wl.qx, wl.qy)
dt = 1e-7
for step in range(1, 50):
    print("STEP", step, flush=True)
    ml = ch.oneTimeStep2D(0.0, dt)
    dt = 0.3 * wl.mesh.min_edge_length / ml
print("SURVIVED", flush=True)
''' % ([ROOT, os.path.join(ROOT, "dg-rain_amd")], N)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    lines = r.stdout.splitlines()
    steps = [int(l.split()[1]) for l in lines if l.startswith("STEP")]
    msg = [l for l in lines if "not a number" in l]
    return steps[-1] if steps else -1, (msg[0] if msg else ""), r.returncode


def main():
    oracle.build(ref=True)
    rng = np.random.default_rng(20261203)
    out = {}

    # 1. Roe flux (numericalFlux2D.c:6-108)
    m = perturbed_mesh(N=6, seed=1)
    ref = oracle.Ref2D(m)
    X = roe_cases(rng)
    F = np.zeros((len(X), 3))
    lam = np.zeros(len(X))
    for i, row in enumerate(X):
        F[i], lam[i] = oracle.ref_roe_flux(*row)
    np.savez_compressed(os.path.join(HERE, "roe_flux.npz"), inputs=X, F=F, lam=lam)

    # 2. one-call operators on a 6x6 mesh with boundary codes 1, 2, 3
    z, qx, qy = random_state(m, seed=11)
    ref.set_state(z, qx, qy)
    r, f, ml = ref.compute2DL()
    SI = np.zeros((m.num_el, 9))
    VI = np.zeros((m.num_el, 21))
    for el in range(m.num_el):
        SI[el], VI[el] = ref.inner_products(el)
    ref.SlopeLimiter()
    lim = np.stack(ref.get_state())
    wd0 = rng.integers(-1, 2, m.num_el).astype(np.int32)
    zs = z - 0.5 * rng.uniform(0, 1, z.size) * (rng.uniform(size=z.size) < 0.3) - 1.9 * (rng.uniform(size=z.size) < 0.1)
    ref.set_state(zs, qx, qy)
    ref.set_wd(wd0)
    ref.wetDryStatus2D()
    wd1 = ref.get_wd()
    ref.PDop2D()
    pd = np.stack(ref.get_state())
    np.savez_compressed(os.path.join(HERE, "operators_6x6.npz"), **mesh_dict(m), zeta=z, qx=qx, qy=qy,
                        rhs=np.stack(r), fhat=np.stack(f), max_lambda=np.array(ml), SI=SI, VI=VI, limited=lim,
                        wd_state=zs, wd_in=wd0, wd_out=wd1, pdop=pd)

    # 3. trajectories through the 2DCode step (no wet/dry, as shipped)
    for name, mm, st, n in (
        ("traj_dam_break_8x8", configs.dam_break(N=8, L=1000.0).mesh, None, 30),
        ("traj_prescribed_6x6", perturbed_mesh(N=6, seed=4), random_state(perturbed_mesh(N=6, seed=4), seed=9, dz=0.05, q=0.05), 25),
        ("traj_lake_at_rest_8x8", configs.lake_at_rest(N=8).mesh, None, 20),
    ):
        if name.startswith("traj_dam"):
            wl = configs.dam_break(N=8, L=1000.0)
            st = (wl.zeta, wl.qx, wl.qy)
        elif name.startswith("traj_lake"):
            wl = configs.lake_at_rest(N=8)
            st = (wl.zeta, wl.qx, wl.qy)
        ref = oracle.Ref2D(mm)
        ref.set_state(*st)
        states, dts = ref_dt_run(ref, n)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **mesh_dict(mm), init=np.stack(st), states=states, dts=dts)

    # 4. wet/dry trajectories through ChanNet -DWDON (walls only)
    for name, mm, st, n in (
        ("traj_wetdry_beach_10x10", beach_mesh(N=10), None, 2),   # step 3 puts NaN into both-dry fluxes, where ChanNet folds with fmax and 2DCode with its max macro
        ("traj_dry_bed_12x12", configs.dry_bed_dam_break(N=12).mesh, None, 40),
    ):
        if "beach" in name:
            zz, qq, rr = beach_state(mm)
            qq = qq * 0.5
            rr = rr * 0.5
            st = (zz, qq, rr)
        else:
            wl = configs.dry_bed_dam_break(N=12)
            st = (wl.zeta, wl.qx, wl.qy)
        ch = oracle.RefChan(mm)
        ch.set_state(*st)
        states, wds, dts = chan_run(ch, n)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **mesh_dict(mm), init=np.stack(st), states=states,
                            wds=wds, dts=dts)

    # 5. topology: the reference's extract_mesh_attributes on a fort.14 file
    mt = perturbed_mesh(N=5, seed=7)
    p14 = os.path.join(HERE, "topology_5x5.14")
    M.write_fort14(p14, mt, title="perturbed 5x5 grid with boundary codes 1 2 3")
    ref = oracle.Ref2D(mt)
    topo = ref.topology_from_fort14(p14)
    np.savez_compressed(os.path.join(HERE, "topology_5x5.npz"), **{k: np.asarray(v) for k, v in topo.items()})

    # 6. the reference's failure on the still-water coastal start
    step, msg, rc = still_coastal_failure()
    np.savez_compressed(os.path.join(HERE, "still_coastal_failure.npz"), step=np.array(step),
                        message=np.array(msg), returncode=np.array(rc))
    print("golden vectors written; still-water coastal: reference stopped in step", step, repr(msg))


if __name__ == "__main__":
    main()
