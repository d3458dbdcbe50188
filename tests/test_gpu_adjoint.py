"""Gradients on the GPU: the hand-written reverse of the RK step (through the
C ABI, dgr_adjoint_gauges) against central finite differences of the forward
solver, for a gauge-misfit loss with respect to nodal depth z and Manning n.

The loss sums (h - obs)^2 over gauge elements sampled every few steps; the
observations come from a run with a perturbed Manning field.  Each check
compares <grad, d> with (L(p + eps d) - L(p - eps d)) / (2 eps) for random
directions d; the tolerance is written in each test."""
import dataclasses

import numpy as np
import pytest

import oracle
from dgrain import configs, native

pytestmark = pytest.mark.gpu


def _problem(N=10, nsteps=30, sample_every=5, ngauge=6, seed=0, rain_mm_h=0.0):
    """Coastal problem started from a generic state: the flood level plus
    small random elevations and flows, so that no limiter bound, node
    extremum or max-|lambda| operand is an exact tie (a central difference
    across an exact tie averages two branches; the reverse takes one)."""
    wl = configs.coastal(N=N, nsteps=nsteps, rain_mm_h=rain_mm_h)
    rng = np.random.default_rng(seed)
    n3 = wl.zeta.size
    wl = dataclasses.replace(wl, zeta=wl.zeta + 0.05 * rng.standard_normal(n3),
                             qx=0.05 * rng.standard_normal(n3), qy=0.05 * rng.standard_normal(n3))
    gauges = np.sort(rng.choice(wl.num_el, ngauge, replace=False)).astype(np.int32)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.set_nodal(n=wl.mesh.n * np.exp(0.2 * rng.standard_normal(wl.mesh.num_nodes)))
    obs = s.gauge_run(nsteps, sample_every, gauges)
    s.close()
    return wl, gauges, obs


def _loss_gpu(wl, gauges, obs, nsteps, sample_every, z=None, n=None):
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_nodal(z=z, n=n)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    h = s.gauge_run(nsteps, sample_every, gauges)
    s.close()
    return float(((h - obs) ** 2).sum())


def _smooth_direction(mesh, rng):
    # a smooth random field over the nodes (a few Fourier modes)
    x = (mesh.x - mesh.x.min()) / (np.ptp(mesh.x) + 1e-300)
    y = (mesh.y - mesh.y.min()) / (np.ptp(mesh.y) + 1e-300)
    d = np.zeros(mesh.num_nodes)
    for _ in range(4):
        kx, ky = rng.integers(0, 3, 2)
        d += rng.standard_normal() * np.cos(np.pi * kx * x + rng.uniform(0, 6.3)) * np.cos(np.pi * ky * y)
    return d


def test_adjoint_loss_matches_forward(gpu):
    nsteps, every = 30, 5
    wl, gauges, obs = _problem(nsteps=nsteps, sample_every=every)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    loss, gz, gn = s.adjoint_gauges(nsteps, every, 7, gauges, obs)
    ref = _loss_gpu(wl, gauges, obs, nsteps, every)
    assert abs(loss[0] - ref) <= 1e-13 * abs(ref), (loss, ref)
    # the handle's state is restored: a second call gives the same bits
    loss2, gz2, gn2 = s.adjoint_gauges(nsteps, every, 7, gauges, obs)
    assert loss2[0] == loss[0] and np.array_equal(gz, gz2) and np.array_equal(gn, gn2)
    assert np.all(np.isfinite(gz)) and np.all(np.isfinite(gn)) and np.abs(gn).max() > 0


def _segment_bytes(s, nn, K, compact=False):
    """Device bytes of one segment's tape (dgr_adjoint_api.hip): four state
    copies and two flag sets per step, plus node bounds and their elements
    unless the tape is compact (the reverse then recomputes them)."""
    _, P = s.device_state()
    up = lambda b: (b + 255) // 256 * 256
    b = 4 * up(9 * P * 8 * K) + 2 * up(P * 4 * K)
    return b if compact else b + up(12 * nn * K * 8) + up(12 * nn * K * 4)


@pytest.mark.parametrize("wetdry", [0, 1])
def test_resident_tapes_give_the_same_gradient(gpu, wetdry):
    """Segments taped by the forward pass and kept resident, segments
    recomputed from checkpoints, and mixes of both (of five segments, the
    last one partial: the last two resident with full tapes; the last three
    with compact tapes, which a budget picks when they keep more segments
    resident, the reverse then recomputing the node bounds) give the same
    loss and gradients bit for bit."""
    from helpers import perturbed_mesh, random_state
    nsteps, every, K = 30, 5, 7
    if wetdry:
        wl, gauges, obs = _problem(nsteps=nsteps, sample_every=every)
        mesh, state, prm = wl.mesh, (wl.zeta, wl.qx, wl.qy), wl.solver_params()
    else:
        mesh = perturbed_mesh(N=10, seed=4)
        state = random_state(mesh, seed=9, dz=0.05, q=0.05)
        prm = native.default_params(cfl=0.3)
        gauges = np.array([3, 40, 77, 150], np.int32)
        s = native.Solver(mesh, prm)
        s.set_state(*state)
        s.set_nodal(n=mesh.n * 1.3)
        obs = s.gauge_run(nsteps, every, gauges)
        s.close()
    s = native.Solver(mesh, prm)
    seg = _segment_bytes(s, mesh.num_nodes, K)
    cseg = _segment_bytes(s, mesh.num_nodes, K, compact=True)
    assert 2 * seg < 3 * cseg <= 2 * seg + seg // 2 < 3 * seg   # full: 2 resident; compact: 3
    out = []
    for budget in (0, 2 * seg, 2 * seg + seg // 2, -1, 1 << 40):
        s.set_tape_budget(budget)
        s.set_state(*state)
        out.append(s.adjoint_gauges(nsteps, every, K, gauges, obs))
    for loss, gz, gn in out[1:]:
        assert loss[0] == out[0][0][0]
        assert np.array_equal(gz, out[0][1]) and np.array_equal(gn, out[0][2])
    assert np.abs(out[0][2]).max() > 0


@pytest.mark.parametrize("still", [True, False])
def test_recorded_ties_equal_counted_ties(gpu, still):
    """The edge-kernel variant records how many max-|lambda| operands equal
    the step's max in the taped forward; the element variants count them in
    the reverse (k_adj_ties).  Same forward bits, so the gradients must agree
    bit for bit, also on still water over a flat bed where every edge ties."""
    from helpers import perturbed_mesh, random_state
    nsteps, every = 12, 3
    if still:
        wl = configs.lake_at_rest(N=8, bumps=False)
        mesh, state = wl.mesh, (wl.zeta, wl.qx, wl.qy)
        prm = native.default_params(**wl.params)
    else:
        mesh = perturbed_mesh(N=10, seed=4)
        state = random_state(mesh, seed=9, dz=0.05, q=0.05)
        prm = native.default_params(cfl=0.3)
    gauges = np.array([3, 40, 77], np.int32)
    s = native.Solver(mesh, prm)
    s.set_state(*state)
    s.set_nodal(n=mesh.n * 1.3 + 0.01)
    obs = s.gauge_run(nsteps, every, gauges) + 1e-3
    out = []
    for variant in (2, 0, 1):
        s.set_variant(variant)
        s.set_state(*state)
        out.append(s.adjoint_gauges(nsteps, every, 4, gauges, obs))
    for loss, gz, gn in out[1:]:
        assert loss[0] == out[0][0][0]
        assert np.array_equal(gz, out[0][1]) and np.array_equal(gn, out[0][2])
    assert np.abs(out[0][1]).max() > 0


@pytest.mark.parametrize("param", ["z", "n"])
def test_adjoint_directional_derivatives(gpu, param):
    """Tolerance: 1e-6 relative between <grad, d> and the central difference,
    in each of 4 random directions; eps relative to the field is 1e-6 for z
    and 1e-4 for n (Manning n ~ 0.025: a smaller step drowns the difference
    in the loss's rounding)."""
    nsteps, every = 40, 5
    wl, gauges, obs = _problem(N=10, nsteps=nsteps, sample_every=every, seed=1, rain_mm_h=20.0)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    loss, gz, gn = s.adjoint_gauges(nsteps, every, 8, gauges, obs)
    s.close()
    g = gz if param == "z" else gn
    base = wl.mesh.z if param == "z" else wl.mesh.n
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(4):
        d = _smooth_direction(wl.mesh, rng) * np.abs(base).mean()
        eps = 1e-6 if param == "z" else 1e-4
        kw_p = {param: base + eps * d}
        kw_m = {param: base - eps * d}
        fd = (_loss_gpu(wl, gauges, obs, nsteps, every, **kw_p) -
              _loss_gpu(wl, gauges, obs, nsteps, every, **kw_m)) / (2 * eps)
        ad = float(g @ d)
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        print("  %s: adjoint %.12e  central difference %.12e  rel %.2e" % (param, ad, fd, rel), flush=True)
        worst = max(worst, rel)
    assert worst <= 1e-6, worst


def test_adjoint_against_oracle_differences(gpu):
    """The gradient against central differences of the CPU oracle (the
    restatement of the reference solver) on the same problem; tolerance 1e-6
    relative, 2 directions for z and 2 for n."""
    nsteps, every = 20, 5
    wl, gauges, obs = _problem(N=8, nsteps=nsteps, sample_every=every, seed=2)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    loss, gz, gn = s.adjoint_gauges(nsteps, every, 5, gauges, obs)
    s.close()

    def oracle_loss(mesh):
        w2 = dataclasses.replace(wl, mesh=mesh)
        o = oracle.Oracle.for_workload(w2)
        hs = []
        for k in range(nsteps // every):
            o.run(every)
            z, _, _ = o.get_state()
            z = z.reshape(-1, 3)
            ev = mesh.elto_vert.reshape(-1, 3)
            zm = 0.5 * (mesh.z[ev] + mesh.z[np.roll(ev, -1, axis=1)])
            hs.append(((z + zm).sum(axis=1) / 3.0)[gauges])
        return float(((np.array(hs)[None] - obs) ** 2).sum())

    rng = np.random.default_rng(9)
    worst = 0.0
    for param, g in (("z", gz), ("n", gn), ("z", gz), ("n", gn)):
        base = getattr(wl.mesh, param)
        d = _smooth_direction(wl.mesh, rng) * np.abs(base).mean()
        eps = 1e-6 if param == "z" else 1e-4
        mp = dataclasses.replace(wl.mesh, **{param: base + eps * d})
        mm = dataclasses.replace(wl.mesh, **{param: base - eps * d})
        fd = (oracle_loss(mp) - oracle_loss(mm)) / (2 * eps)
        ad = float(g @ d)
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        print("  %s: adjoint %.12e  oracle central difference %.12e  rel %.2e" % (param, ad, fd, rel), flush=True)
        worst = max(worst, rel)
    assert worst <= 1e-6, worst


def _dry_bed_problem(N=24, W=40, seed=5):
    """The dry-bed dam break (5 mm reservoir for x <= 5 m, dry film beyond,
    walls) made tie-free so that the loss is differentiable where it is
    evaluated.  Three exact ties of the plain workload are removed:
    * the dry film sits at 0.8 H0 instead of exactly H0, PDop2D's own
      threshold (wetDry2D.c:93: Ht <= H0);
    * the reservoir surface carries a smooth 0.2 mm perturbation (still water
      of uniform depth makes every wet edge tie for max |lambda|, the dt
      operand);
    * the bed carries a smooth 20 um relief.
    Manning n is 0.002: with n = 0.02 the explicit friction of the reference
    scheme diverges in the thin films (its NaN failure) within a few steps.
    Every element is a gauge, sampled after each step; the observations come
    from a run with n scaled by 1.2."""
    from dgrain import mesh as M
    from helpers import smooth_field
    wl = configs.dry_bed_dam_break(N=N, nsteps=W)
    rng = np.random.default_rng(seed)
    mesh = dataclasses.replace(wl.mesh, z=wl.mesh.z + 2e-5 * smooth_field(wl.mesh.x, wl.mesh.y, rng),
                               n=np.full(wl.mesh.num_nodes, 0.002))
    surf = np.where(mesh.x <= 5.0, 0.005 + 2e-4 * smooth_field(mesh.x, mesh.y, rng), 0.8e-3) - mesh.z
    wl = dataclasses.replace(wl, mesh=mesh, zeta=M.nodal_to_midpoints(mesh, surf))
    gauges = np.arange(wl.num_el, dtype=np.int32)
    obs = _gauges_of(wl, gauges, W, n=mesh.n * 1.2)
    return wl, gauges, obs


def _gauges_of(wl, gauges, W, z=None, n=None):
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_nodal(z=z, n=n)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    h = s.gauge_run(W, 1, gauges)
    s.close()
    return h


def test_adjoint_through_pdop(gpu):
    """The gradient through PDop2D's mass and momentum corrections
    (wetDry2D.c:93-253), differentiated on the branch each element took:
    40 steps of the tie-free dry-bed dam break (24x24 cells), in which the
    oracle counts thousands of element-stages that PDop changed.  dL/dn in 8
    smooth random directions and dL/dz in 8 smooth directions over the
    reservoir, against settled central differences of the GPU forward;
    tolerance 1e-6.

    A bed perturbation under the dry film is left out on purpose: PDop pins
    a lifted vertex at exactly H0, and the next stage's PDop switches its
    momentum redistribution on a vertex depth <= H0 test, so moving the bed
    under a pinned vertex jumps the state (the reference scheme is
    discontinuous there, and no central difference settles)."""
    from helpers import converged_difference, smooth_field
    W = 40
    wl, gauges, obs = _dry_bed_problem(W=W)
    o = oracle.Oracle.for_workload(wl)
    o.run(W)
    changed = o.pdop_changed
    o.close()
    assert changed > 1000, changed
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    loss, gz, gn = s.adjoint_gauges(W, 1, 8, gauges, obs)
    s.close()
    ref = float(((_gauges_of(wl, gauges, W) - obs) ** 2).sum())
    assert abs(loss[0] - ref) <= 1e-13 * ref
    print("  PDop changed %d element-stages; loss %.6e" % (changed, loss[0]), flush=True)
    rng = np.random.default_rng(9)
    m = wl.mesh
    worst = 0.0
    for param, g in (("n", gn), ("z", gz)):
        base = m.n if param == "n" else m.z
        for j in range(8):
            d = smooth_field(m.x, m.y, rng)
            d = d * np.abs(base).mean() if param == "n" else d * np.clip((3.5 - m.x) / 1.5, 0.0, 1.0)
            eps = [1e-5, 1e-6, 1e-7] if param == "n" else [1e-7, 1e-8, 1e-9, 1e-10]

            def L(e):
                kw = {param: base + e * d}
                return float(((_gauges_of(wl, gauges, W, **kw) - obs) ** 2).sum())

            fd, e = converged_difference(L, eps, 1e-6)
            assert fd is not None, "%s direction %d: central differences did not settle" % (param, j)
            ad = float(g @ d)
            rel = abs(ad - fd) / abs(fd)
            print("  %s dir %d: adjoint %.12e  central difference %.12e (eps %.0e)  rel %.2e" % (param, j, ad, fd, e, rel),
                  flush=True)
            worst = max(worst, rel)
    assert worst <= 1e-6, worst


def test_adjoint_runs_on_the_plain_dry_bed(gpu):
    """The plain dry-bed dam break (film at exactly H0): PDop acts on every
    dry element each stage and the adjoint returns a gradient (it used to
    refuse such runs); its loss is the forward's.  Central differences do
    not apply here: the film sits on PDop's threshold, so the loss has kinks
    at the evaluation point."""
    W = 40
    wl = configs.dry_bed_dam_break(N=24, nsteps=W)
    gauges = np.arange(0, wl.num_el, 7, dtype=np.int32)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    obs = s.gauge_run(W, 5, gauges) + 1e-3    # (a larger n makes the thin films' friction diverge)
    s.close()
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    loss, gz, gn = s.adjoint_gauges(W, 5, 10, gauges, obs)
    s.close()
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    ref = float(((s.gauge_run(W, 5, gauges) - obs) ** 2).sum())
    s.close()
    assert abs(loss[0] - ref) <= 1e-13 * ref
    assert np.all(np.isfinite(gz)) and np.all(np.isfinite(gn)) and np.abs(gn).max() > 0


def test_autograd_harness(gpu):
    """torch.autograd over the C ABI: backward gives the adjoint gradients,
    and a few gradient steps on Manning n reduce the misfit."""
    import torch
    from dgrain import autograd
    nsteps, every = 30, 5
    wl, gauges, obs = _problem(N=10, nsteps=nsteps, sample_every=every, seed=3)
    prob = autograd.GaugeProblem(wl.mesh, wl.solver_params(), wl.zeta, wl.qx, wl.qy, gauges, obs[0], nsteps,
                                 every, 6)
    z = torch.tensor(wl.mesh.z, dtype=torch.float64, requires_grad=True)
    n = torch.tensor(wl.mesh.n, dtype=torch.float64, requires_grad=True)
    L = autograd.gauge_loss(prob, z, n)
    L.backward()
    ref_loss, gz, gn = prob.loss_and_grads(wl.mesh.z, wl.mesh.n)
    assert float(L) == ref_loss
    assert np.array_equal(z.grad.numpy(), gz) and np.array_equal(n.grad.numpy(), gn)
    # the misfit of the forward-only path is the same number
    assert abs(prob.loss(wl.mesh.z, wl.mesh.n) - ref_loss) <= 1e-13 * ref_loss
    # steepest descent on log n with a backtracking step
    logn = torch.log(torch.tensor(wl.mesh.n, dtype=torch.float64)).requires_grad_(True)
    zt = torch.tensor(wl.mesh.z, dtype=torch.float64)
    losses = []
    step = 1.0
    for it in range(4):
        L = autograd.gauge_loss(prob, zt, torch.exp(logn))
        L.backward()
        losses.append(float(L))
        with torch.no_grad():
            g = logn.grad
            while True:
                trial = logn - step * g / (g.abs().max() + 1e-300) * 0.05
                if prob.loss(wl.mesh.z, torch.exp(trial).numpy()) < float(L) or step < 1e-3:
                    break
                step *= 0.5
            logn.copy_(trial)
            logn.grad.zero_()
    print("  misfit by iteration:", ["%.4e" % v for v in losses])
    assert losses[-1] < losses[0]
