"""Gradients at the sizes of the inversion workloads.

Manning-n inversion (131,072 triangles: the 256x256 coastal mesh, bed,
n field and tide of the inundation workload; 64 gauges sampled every 10
steps; the starting n the true field times a lognormal factor, sigma 0.2):
the adjoint's dL/dz and dL/dn against central differences of the GPU
forward in 16 smooth random directions each, and against central
differences of the oracle in its reference-math mode (the arithmetic of the
compiled reference) in one direction each, with the tapes recomputed from
checkpoints and with the tapes kept resident (the same gradient bits).

Why a smooth start and a short window.  The discrete step is piecewise
smooth: limiter clamps, the max-|lambda| argmax behind dt and the entropy
fix are kinks.  A central difference across a kink averages two branches,
while the adjoint differentiates the branch the run takes.  Over 131k
elements the kinks are dense: from still water (exact ties everywhere) or
a per-midpoint random state (the limiter clips almost everywhere) central
differences do not settle at any step size, and from a smooth start they
settle only over the first 50-100 steps.  Each difference is therefore
taken at the largest step size at which it agrees with the next smaller
one (no kink crossed), and averaged over a few nearby step sizes to damp
the rounding noise of separate runs.  dL/dn is weaker still: over the 20
steps friction moves the gauges by ~1e-3 m, and the trajectory's own
rounding (~1e-14 m at the gauges) puts the noise of a single central
difference at ~3e-6 relative; its tolerance is 1e-5.

Batched gradients (several members of uneven size, one 512x512, each with
its own 64 gauges): every member's loss and gradients equal the same
member's solo gradient bit for bit, recomputing and resident, and one
member's gradient matches central differences."""
import dataclasses
import os

import numpy as np
import pytest

import oracle
from dgrain import configs, native
from helpers import converged_difference, smooth_field, smooth_start

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _problem(N, W, seed, amp, qamp, every=10, ngauge=64, cfg_seed=2):
    wl = configs.coastal(N=N, nsteps=W, seed=cfg_seed)
    rng = np.random.default_rng(seed)
    wl = smooth_start(wl, rng, amp, qamp)
    gauges = np.sort(rng.choice(wl.num_el, ngauge, replace=False)).astype(np.int32)
    truth = wl.mesh.n.copy()
    start = truth * np.exp(0.2 * rng.standard_normal(truth.size))
    obs = _gauges_gpu(wl, gauges, W, every, wl.mesh.z, truth)
    return wl, gauges, start, obs


def _gauges_gpu(wl, gauges, W, every, z, n):
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_nodal(z=z, n=n)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    h = s.gauge_run(W, every, gauges)
    s.close()
    return h


def _gauges_oracle(wl, gauges, W, every, z, n):
    mesh = dataclasses.replace(wl.mesh, z=z, n=n)
    o = oracle.Oracle(mesh, math=oracle.MATH_REF, threads=THREADS, **wl.params)
    o.set_state(wl.zeta, wl.qx, wl.qy)
    ev = mesh.elto_vert.reshape(-1, 3)
    zm = 0.5 * (z[ev] + z[np.roll(ev, -1, axis=1)])
    hs = []
    for _ in range(W // every):
        o.run(every)
        zeta = o.get_state()[0].reshape(-1, 3)
        hs.append(((zeta + zm).sum(axis=1) / 3.0)[gauges])
    o.close()
    return np.array(hs)[None]


def _adjoint(wl, gauges, start, obs, W, every, K, budget):
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_nodal(n=start)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.set_tape_budget(budget)
    out = s.adjoint_gauges(W, every, K, gauges, obs)
    s.close()
    return out


def _check_directions(wl, gauges, start, obs, W, every, grad, param, ndir, eps_list, tol, rng, runner, avg=1):
    base = start if param == "n" else wl.mesh.z
    worst = 0.0
    for j in range(ndir):
        d = smooth_field(wl.mesh.x, wl.mesh.y, rng) * np.abs(base).mean()

        def L(e):
            kw = dict(z=wl.mesh.z, n=start)
            kw[param] = base + e * d
            return float(((runner(wl, gauges, W, every, kw["z"], kw["n"]) - obs) ** 2).sum())

        fd, e = converged_difference(L, eps_list, tol, avg)
        assert fd is not None, "%s direction %d: central differences did not settle" % (param, j)
        ad = float(grad @ d)
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        print("  %s dir %2d: adjoint %.12e  central difference %.12e (eps %.1e)  rel %.2e" % (param, j, ad, fd, e, rel),
              flush=True)
        worst = max(worst, rel)
    assert worst <= tol, worst
    return worst


def test_inversion_256_dz(gpu):
    """dL/dz over 50 steps (5 samples, checkpoints every 10): 16 directions
    against the GPU forward, 1 against the reference-math oracle; 1e-6."""
    W, every = 50, 10
    wl, gauges, start, obs = _problem(256, W, seed=11, amp=0.05, qamp=0.05)
    loss, gz, gn = _adjoint(wl, gauges, start, obs, W, every, 10, 0)
    loss2, gz2, gn2 = _adjoint(wl, gauges, start, obs, W, every, 10, -1)
    assert loss2[0] == loss[0] and np.array_equal(gz, gz2) and np.array_equal(gn, gn2)
    ref = float(((_gauges_gpu(wl, gauges, W, every, wl.mesh.z, start) - obs) ** 2).sum())
    assert abs(loss[0] - ref) <= 1e-13 * ref
    rng = np.random.default_rng(21)
    eps = [4e-7, 1e-7, 2.5e-8, 6.25e-9]
    _check_directions(wl, gauges, start, obs, W, every, gz, "z", 16, eps, 1e-6, rng, _gauges_gpu, avg=2)
    _check_directions(wl, gauges, start, obs, W, every, gz, "z", 1, eps, 1e-6, rng, _gauges_oracle)


def test_inversion_256_dn(gpu):
    """dL/dn over 20 steps (2 samples, checkpoints every 5) with a stronger
    smooth flow (discharge amplitude 1 m^2/s, so friction moves the gauges
    well above rounding): 16 directions against the GPU forward, 1 against
    the reference-math oracle; 1e-5 (the rounding floor of the difference)."""
    W, every = 20, 10
    wl, gauges, start, obs = _problem(256, W, seed=11, amp=0.05, qamp=1.0)
    loss, gz, gn = _adjoint(wl, gauges, start, obs, W, every, 5, 0)
    loss2, gz2, gn2 = _adjoint(wl, gauges, start, obs, W, every, 5, -1)
    assert loss2[0] == loss[0] and np.array_equal(gz, gz2) and np.array_equal(gn, gn2)
    rng = np.random.default_rng(22)
    eps = [2e-6, 1e-6, 5e-7]
    _check_directions(wl, gauges, start, obs, W, every, gn, "n", 16, eps, 1e-5, rng, _gauges_gpu, avg=4)
    _check_directions(wl, gauges, start, obs, W, every, gn, "n", 1, eps, 1e-5, rng, _gauges_oracle, avg=4)


def test_batched_gradients_equal_solo(gpu):
    """Four members of uneven size (512x512, 97x97, 200x200, 33x33 cells),
    each with its own bumps, n field, tide amplitude, smooth start, 64 gauges
    and observations, in one ensemble handle: losses and gradients bit for
    bit the members' solo gradients, with every segment recomputed and with
    resident tapes; then member 1's dL/dz against central differences of its
    solo forward in 2 directions (1e-6)."""
    W, every, K = 40, 10, 10
    sizes = (512, 97, 200, 33)
    probs = []
    for i, N in enumerate(sizes):
        wl = configs.coastal(N=N, L=N * 1.0e4 / 512, nsteps=W, seed=40 + i, tide_amp=0.5 + 0.5 * i, level=3.5)
        rng = np.random.default_rng(60 + i)
        wl = smooth_start(wl, rng)
        gauges = np.sort(rng.choice(wl.num_el, 64, replace=False)).astype(np.int32)
        truth = wl.mesh.n.copy()
        start = truth * np.exp(0.2 * rng.standard_normal(truth.size))
        obs = _gauges_gpu(wl, gauges, W, every, wl.mesh.z, truth)
        probs.append((wl, gauges, start, obs))
    solo = {b: [_adjoint(wl, g, s, o, W, every, K, b) for wl, g, s, o in probs] for b in (0, -1)}
    for b in (0, -1):
        e = native.Ensemble([p[0].mesh for p in probs], [p[0].solver_params() for p in probs])
        for i, (wl, g, s, o) in enumerate(probs):
            e.set_nodal(n=s, member=i)
            e.set_state(i, wl.zeta, wl.qx, wl.qy)
        e.set_tape_budget(b)
        loss, gz, gn = e.adjoint_gauges(W, every, K, np.stack([p[1] for p in probs]),
                                        np.concatenate([p[3] for p in probs]))
        e.close()
        off = 0
        for i, (wl, g, s, o) in enumerate(probs):
            nn = wl.mesh.num_nodes
            sl, sz, sn = solo[b][i]
            assert loss[i] == sl[0], (b, i)
            assert np.array_equal(gz[off:off + nn], sz) and np.array_equal(gn[off:off + nn], sn), (b, i)
            off += nn
        print("  budget %d: %d members bit for bit their solo gradients" % (b, len(probs)), flush=True)
    for i in range(len(probs)):
        assert np.array_equal(solo[0][i][1], solo[-1][i][1]) and np.array_equal(solo[0][i][2], solo[-1][i][2])
    wl, g, s, o = probs[1]
    rng = np.random.default_rng(23)
    _check_directions(wl, g, s, o, W, every, solo[0][1][1], "z", 2, [4e-7, 1e-7, 2.5e-8, 6.25e-9], 1e-6, rng,
                      _gauges_gpu, avg=2)


# padded edges from which the library reverses wet/dry edge fluxes in two
# launches (wet edges, then the listed others), when at most 1 in 1000
# elements is not wet as the reverse starts
EDGE_SPLIT = 1 << 21


def _padded_edges(wl):
    ned = int(np.asarray(wl.mesh.edg_to_vert).size // 2)
    return (ned + 255) // 256 * 256


def test_split_edge_reverse_equals_one_pass(gpu):
    """Three 512x512 coastal members (all wet, own bumps, tide and gauges)
    and a 32x32 dry-bed dam break (half its domain dry, its other edges one-
    or both-side dry as the front moves) in one ensemble: 2.36 M padded edges
    and 0.07% of the elements not wet, so the edge reverse runs as a wet-edge
    launch plus a launch over the listed non-wet edges.  Every member's loss
    and gradients equal its solo gradient (one launch per edge reverse) bit
    for bit."""
    W, every, K = 20, 10, 10
    probs = []
    for i in range(4):
        if i < 3:
            wl = configs.coastal(N=512, nsteps=W, seed=70 + i, tide_amp=0.5 + 0.5 * i, level=3.5)
            rng = np.random.default_rng(80 + i)
            wl = smooth_start(wl, rng)
        else:
            # Manning n 0.002: with 0.02 the reference scheme's explicit
            # friction diverges in the thin films of this small mesh
            wl = configs.dry_bed_dam_break(N=32, nsteps=W)
            wl = dataclasses.replace(wl, mesh=dataclasses.replace(wl.mesh, n=np.full(wl.mesh.num_nodes, 0.002)))
            rng = np.random.default_rng(83)
        gauges = np.sort(rng.choice(wl.num_el, 64, replace=False)).astype(np.int32)
        truth = wl.mesh.n.copy()
        start = truth * np.exp(0.2 * rng.standard_normal(truth.size))
        obs = _gauges_gpu(wl, gauges, W, every, wl.mesh.z, truth)
        probs.append((wl, gauges, start, obs))
    assert all(_padded_edges(p[0]) < EDGE_SPLIT for p in probs)
    assert sum(_padded_edges(p[0]) for p in probs) >= EDGE_SPLIT
    solo = [_adjoint(wl, g, s, o, W, every, K, -1) for wl, g, s, o in probs]
    e = native.Ensemble([p[0].mesh for p in probs], [p[0].solver_params() for p in probs])
    for i, (wl, g, s, o) in enumerate(probs):
        e.set_nodal(n=s, member=i)
        e.set_state(i, wl.zeta, wl.qx, wl.qy)
    loss, gz, gn = e.adjoint_gauges(W, every, K, np.stack([p[1] for p in probs]),
                                    np.concatenate([p[3] for p in probs]))
    e.close()
    off = 0
    for i, (wl, g, s, o) in enumerate(probs):
        nn = wl.mesh.num_nodes
        sl, sz, sn = solo[i]
        assert np.all(np.isfinite(sz)) and np.any(sz != 0), i
        assert loss[i] == sl[0], i
        assert np.array_equal(gz[off:off + nn], sz) and np.array_equal(gn[off:off + nn], sn), i
        off += nn
    print("  %d members of a split edge reverse bit for bit their one-launch gradients" % len(probs), flush=True)
