"""Benchmark: FP64 DOF-updates/s per RK stage of the 2D DG shallow-water step.

Workload (one MI355X): the coastal inundation mesh of 512 x 512 cells split
into 524,288 triangles over 10 km x 10 km, bed sloping from -10 m to +2 m plus
32 seeded Gaussian bumps, lognormal Manning n (median 0.025, sigma 0.3), tide
1.5 m sin(2 pi t / 43200 s) on x = 0, reflecting walls elsewhere, p=1, wet/dry
and limiter on, SSP-RK2 at CFL 0.3, started fully wet at +2.5 m (the
still-water start stops on a NaN flux in the reference itself within 3 steps).

A step is one SSP-RK2 step (two RK stages) of the whole mesh; a DOF-update is
one of the 9 unknowns per element advanced through one stage.  Inputs are
resident in HBM before timing; dt is computed on the device.

Prints one JSON line (see the "value"/"roofline"/"cpu_baseline" fields).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "dg-rain_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
# FP64 VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction
VALU_PEAK_GWIS = 256 * 4 * 2.4 / 4
# algorithmic HBM bytes per element for one launch of the flux/RHS/update
# kernel, averaged over the two stages: state in 72 + state out 72 + step-start
# state 36 (72 in stage 2 only) + mesh 40 + element means out 24
KSTAGE_BYTES_PER_EL = 244.0
# algorithmic bytes per element per RK stage for the whole stage
STAGE_BYTES_PER_EL = 268.0
# FP64 vector peak of MI355X (FLOP/s)
FP64_PEAK_TFLOPS = 78.6
# the kernel(s) each stage-kernel variant launches for the flux/RHS/update
STAGE_KERNELS = {0: "k_stage (fused flux gather + RHS + RK update)",
                 1: "k_flux + k_rhs (element-centric fluxes, then RHS + RK update)",
                 2: "k_edge_flux + k_rhs (edge-centric fluxes, then RHS + RK update)",
                 3: "k_stage_persist (edge fluxes, RHS + RK update, node bounds, limiter in one cooperative launch)"}
STAGE_ALL = {0: "k_stage, k_node_minmax, k_limit; the step end rides on the stage-2 node-bounds launch",
             1: "k_flux, k_rhs, k_node_minmax, k_limit; the step end rides on the stage-2 node-bounds launch",
             2: "k_edge_flux, k_rhs, k_node_minmax, k_limit; the step end rides on the stage-2 node-bounds launch",
             3: "k_stage_persist: edge fluxes, RHS, node bounds, limiter between grid barriers"}


def progress(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--N", type=int, default=512, help="grid cells per side (2*N*N triangles)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--no-wetdry", action="store_true", help="skip the wet/dry-active (dry-bed dam break) leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-cost measurement")
    ap.add_argument("--adj-N", type=int, default=256)
    ap.add_argument("--no-large", action="store_true", help="skip the 2048x2048 run and the batched adjoints")
    ap.add_argument("--large-steps", type=int, default=20)
    ap.add_argument("--batch-adj-steps", type=int, default=100)
    ap.add_argument("--adj-steps", type=int, default=2000)
    ap.add_argument("--ens-members", type=int, default=64)
    ap.add_argument("--ens-steps", type=int, default=20, help="timed steps of the ensemble and of each solo run")
    ap.add_argument("--variant", type=int, default=-1,
                    help="stage-kernel variant to time (default: pick the fastest with dgr_autotune)")
    ap.add_argument("--block", type=int, default=0,
                    help="workgroup size of the element/node kernels (default: pick with dgr_autotune_block)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 512 MiB (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy of known size: the FETCH_SIZE/WRITE_SIZE
    counters of this kernel give the scale between counter units and bytes."""
    import torch
    n = 64 * 2 ** 20
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "dispatches": 5}))


PAIR = {"k_edge_flux<true>": 2, "k_rhs<1>": 1, "k_rhs<2>": 1}   # launches per step of the variant-2 pair


def load_pair(variant):
    """(HBM bytes, VALU wave-instructions) per stage launch of the
    flux/RHS/update kernel pair, from the counter summary in profiles/."""
    c = load_counters(variant)
    if c is None or variant != 2:
        return None, None
    kk = c["kernels"]
    return (sum(kk[k]["bytes"] * n for k, n in PAIR.items()) / 2.0,
            sum(kk[k]["valu_insts"] * n for k, n in PAIR.items()) / 2.0)


def load_counters(variant):
    """Per RK stage of the whole step (all kernels, the step end included):
    HBM bytes from FETCH_SIZE/WRITE_SIZE, wave-level VALU
    instructions from SQ_INSTS_VALU, FP64 FLOPs from SQ_INSTS_VALU_FLOPS_FP64
    (+ _TRANS), from the PMC passes summarised in profiles/ (None when they
    cover another variant)."""
    p = os.path.join(ROOT, "profiles", "coastal_512_stage_counters.json")
    if os.path.exists(p):
        with open(p) as f:
            d = json.load(f)
        if d.get("variant") == variant:
            return d
    return None


def timed_run(s, steps):
    """s.run(steps) timed by the wall clock and by two HIP events recorded on
    the handle's own stream (every kernel of the run is on it); no per-kernel
    events inside the loop."""
    import torch
    st = torch.cuda.ExternalStream(s.stream())
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(st)
    rep = s.run(steps)
    e1.record(st)
    e1.synchronize()
    wall = time.perf_counter() - t0
    return rep, wall, e0.elapsed_time(e1) * 1e-3


def wetdry_counts(mesh, wd):
    """(dry elements, wet/dry edges: interior edges with one side flagged dry
    and the other wet)."""
    import numpy as np
    e2e = np.asarray(mesh.edg_to_els).reshape(-1, 2)
    a, b = wd[e2e[:, 0]], wd[e2e[:, 1]]
    return int((wd == 0).sum()), int(((a == 0) != (b == 0)).sum())


def wetdry_bench(a, variant):
    """A wet/dry-active leg at the headline size: the dam break onto a dry bed
    (N x N cells over 10 m x 10 m, 5 mm reservoir on x <= 5 m, dry film at H0
    elsewhere, walls, wet/dry and limiter on, CFL 0.3), where the wet/dry
    status, the dry-edge Roe fluxes and PDop act every stage.  Timed like the
    headline; the dry elements and wet/dry edges are counted at the start and
    the end of the timed steps."""
    from dgrain import configs, native
    wl = configs.dry_bed_dam_break(N=a.N, nsteps=a.steps)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_variant(variant)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.run(a.warmup)
    s.sync()
    start = wetdry_counts(wl.mesh, s.get_wd())
    rep, wall, ev = timed_run(s, a.steps)
    end = wetdry_counts(wl.mesh, s.get_wd())
    s.close()
    ne = wl.num_el
    return {"workload": "dry_bed_dam_break_%dx%d" % (a.N, a.N), "triangles": ne, "steps": a.steps,
            "value": 9.0 * ne * 2 * a.steps / ev, "unit": "DOF-updates/s", "ms_per_step": ev / a.steps * 1e3,
            "wall_ms_per_step": wall / a.steps * 1e3, "sim_time_s": rep.t,
            "dry_elements": {"start": start[0], "end": end[0]},
            "wet_dry_edges": {"start": start[1], "end": end[1]},
            "stage_roofline_frac": STAGE_BYTES_PER_EL * ne * 2 * a.steps / ev / 1e9 / HBM_PEAK_GBS}


def cpu_baseline(wl, budget_s, threads=1, ref_flags=False):
    """The oracle (C restatement) timed on a bounded number of steps of the
    same workload: device-math mode at -O2, one thread (the reference is
    single-threaded) or `threads` OpenMP threads over the element and edge
    loops (same bits); or (ref_flags) reference-math mode built at the
    reference's own flags, -g with no optimisation (2DCode/Makefile:8)."""
    import oracle
    if ref_flags:
        o = oracle.Oracle.for_workload(wl, math=oracle.MATH_REF, threads=threads, lib=oracle.LIB_G)
    else:
        o = oracle.Oracle.for_workload(wl, math=oracle.MATH_DEV, threads=threads)
    o.run(1)                     # first step uses dt0 = 1e-7; not timed
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.run(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 50:
            break
    dofs = 9.0 * wl.num_el * 2 * steps
    return {"value": dofs / el, "unit": "DOF-updates/s per RK stage", "cores": threads, "kind": "port",
            "sample": "%d SSP-RK2 steps of the same %d-triangle workload, oracle C restatement (%s), %d thread%s, %.1f s"
                      % (steps, wl.num_el, "reference math, -g -O0 as 2DCode/Makefile:8" if ref_flags
                         else "device math, -O2", threads, "" if threads == 1 else "s", el)}


def ensemble_bench(a, variant):
    """The batched-ensemble workload (64 coastal members, N log-uniform in
    [64, 512], own tide amplitude, rain, bumps and n field each): every
    member advanced by one launch per kernel, steps replayed from a captured
    hipGraph, against the plain one-instance-per-launch path (each member on
    its own handle, launches enqueued per step) over the same steps."""
    from dgrain import configs, native
    wls = configs.ensemble(n=a.ens_members, seed=3)
    ne_tot = sum(w.num_el for w in wls)
    ens = native.Ensemble([w.mesh for w in wls], [w.solver_params() for w in wls])
    for i, w in enumerate(wls):
        ens.set_state(i, w.zeta, w.qx, w.qy)
    ens.set_graph(True)
    ens.run(2)
    ens.sync()
    ens_variant, ens_tune = ens.autotune(3)
    ens_block, ens_block_ms = ens.autotune_block(3)
    ens.run(1)
    ens.sync()
    t0 = time.perf_counter()
    ens.run(a.ens_steps)
    ens.sync()
    t_ens = time.perf_counter() - t0
    ens.close()
    t_solo = 0.0
    for w in wls:
        s = native.Solver(w.mesh, w.solver_params())
        s.set_variant(ens_variant)
        s.set_block(ens_block)
        s.set_state(w.zeta, w.qx, w.qy)
        s.run(3)
        s.sync()
        t0 = time.perf_counter()
        s.run(a.ens_steps)
        s.sync()
        t_solo += time.perf_counter() - t0
        s.close()
    dofs = 9.0 * ne_tot * 2 * a.ens_steps
    v_ens, v_solo = dofs / t_ens, dofs / t_solo
    return {"members": len(wls), "triangles": ne_tot, "sizes_N": [int(round((w.num_el / 2) ** 0.5)) for w in wls],
            "steps": a.ens_steps, "variant": ens_variant, "autotune_ms_per_step": ens_tune,
            "block": ens_block, "block_autotune_ms_per_step_64_128_256": ens_block_ms,
            "value": v_ens, "unit": "DOF-updates/s", "ms_per_step": t_ens / a.ens_steps * 1e3,
            "solo_value": v_solo, "solo_ms_per_step": t_solo / a.ens_steps * 1e3,
            "speedup": v_ens / v_solo,
            "stage_roofline_frac": STAGE_BYTES_PER_EL * ne_tot * 2 * a.ens_steps / t_ens / 1e9 / HBM_PEAK_GBS}


def adjoint_bench(a, variant):
    """Adjoint Manning-n inversion workload: the coastal physics on an N x N
    grid (256: 131,072 triangles), a gauge misfit over 64 seeded gauge
    elements sampled every 10 steps, observations from the seeded 'true' n
    field, the gradient taken at a perturbed n (truth times a lognormal
    factor, sigma 0.2) with checkpoints every 50 steps.  Cost = time of one
    dgr_adjoint_gauges call over the time of the plain forward run of the
    same steps, twice: with the default tape budget (the free HBM less 1/16
    of the device memory: the forward tapes every segment that fits and those are reversed without
    recomputation) and with budget 0 (every segment recomputed from its
    checkpoint).  Both give the same gradient bit for bit.  Each is timed on
    a second call, as in a gradient loop: the first call also allocates the
    tape space the handle keeps (reported as adjoint_first_call_s)."""
    import numpy as np
    from dgrain import configs, native
    wl = configs.coastal(N=a.adj_N, start="flood")
    rng = np.random.default_rng(4)
    gauges = np.sort(rng.choice(wl.num_el, 64, replace=False)).astype(np.int32)
    n_true = wl.mesh.n
    n0 = n_true * np.exp(0.2 * rng.standard_normal(n_true.size))
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_variant(variant)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    obs = s.gauge_run(a.adj_steps, 10, gauges)
    s.close()
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_variant(variant)
    s.set_nodal(n=n0)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.adjoint_gauges(50, 10, 50, gauges, obs[:, :5])   # warm the code objects
    s.sync()

    def timed():
        t0 = time.perf_counter()
        r = s.adjoint_gauges(a.adj_steps, 10, 50, gauges, obs)
        s.sync()
        return time.perf_counter() - t0, r

    # the first call allocates the tape space the handle keeps; a gradient
    # loop's steady cost is the second call's
    t_first, _ = timed()
    t_adj, (loss, gz, gn) = timed()
    progress("adjoint (resident tapes) %.3f s, first call %.3f s" % (t_adj, t_first))
    s.set_tape_budget(0)
    timed()
    t_rec, (loss_r, gz_r, gn_r) = timed()
    progress("adjoint (recompute) %.3f s" % t_rec)
    same = bool(loss_r[0] == loss[0] and np.array_equal(gz_r, gz) and np.array_equal(gn_r, gn))
    t0 = time.perf_counter()
    h = s.gauge_run(a.adj_steps, 10, gauges)
    s.sync()
    t_fwd = time.perf_counter() - t0
    s.close()
    return {"workload": "coastal_flood_%dx%d_manning_inversion" % (a.adj_N, a.adj_N), "triangles": wl.num_el,
            "steps": a.adj_steps, "gauges": 64, "sample_every": 10, "checkpoint_every": 50,
            "loss": float(loss[0]), "loss_forward_only": float(((h - obs) ** 2).sum()),
            "grad_n_norm": float(np.linalg.norm(gn)), "grad_z_norm": float(np.linalg.norm(gz)),
            "forward_s": t_fwd, "adjoint_s": t_adj, "ratio": t_adj / t_fwd, "adjoint_first_call_s": t_first,
            "adjoint_recompute_s": t_rec, "ratio_recompute": t_rec / t_fwd, "same_gradient_bits": same}


def large_bench(a, variant):
    """The large workload: part 1, the 2048 x 2048 mesh (8,388,608 triangles, 40 km,
    128 bumps, tide + 20 mm/h rain) timed over a few steps; part 2, 16 gauge
    gradients on one 512 x 512 sub-mesh (64 seeded gauges each), batched as
    the 16 members of one ensemble handle (one launch per reverse kernel for
    all 16), against the forward run of the same batch."""
    import numpy as np
    from dgrain import configs, native
    wl = configs.large()
    progress("large mesh built (%d triangles)" % wl.num_el)
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_variant(variant)
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.run(3)
    s.sync()
    t0 = time.perf_counter()
    s.run(a.large_steps)
    s.sync()
    el = time.perf_counter() - t0
    s.close()
    ne = wl.num_el
    part1 = {"workload": wl.name, "triangles": ne, "steps": a.large_steps,
             "value": 9.0 * ne * 2 * a.large_steps / el, "unit": "DOF-updates/s",
             "ms_per_step": el / a.large_steps * 1e3,
             "stage_roofline_frac": STAGE_BYTES_PER_EL * ne * 2 * a.large_steps / el / 1e9 / HBM_PEAK_GBS}
    del wl
    sub = configs.coastal(N=512, start="flood")
    nb = 16
    rng = np.random.default_rng(6)
    gauges = np.stack([np.sort(rng.choice(sub.num_el, 64, replace=False)) for _ in range(nb)]).astype(np.int32)
    ens = native.Ensemble([sub.mesh] * nb, [sub.solver_params() for _ in range(nb)])
    progress("batched-adjoint ensemble created")
    ens.set_variant(variant)
    for i in range(nb):
        ens.set_state(i, sub.zeta, sub.qx, sub.qy)
    S = a.batch_adj_steps
    obs = ens.gauge_run(S, 10, gauges) * (1.0 + 1e-4)
    for i in range(nb):
        ens.set_state(i, sub.zeta, sub.qx, sub.qy)
    # timed on the second call, as in a gradient loop (the first also
    # allocates the tape space the handle keeps)
    t0 = time.perf_counter()
    ens.adjoint_gauges(S, 10, 50, gauges, obs)
    ens.sync()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss, gz, gn = ens.adjoint_gauges(S, 10, 50, gauges, obs)
    ens.sync()
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    ens.gauge_run(S, 10, gauges)
    ens.sync()
    t_fwd = time.perf_counter() - t0
    ens.close()
    part2 = {"workload": "16 gradients, coastal_flood_512x512 sub-mesh, 64 gauges each", "members": nb,
             "triangles": nb * sub.num_el, "steps": S, "checkpoint_every": 50, "forward_s": t_fwd,
             "adjoint_s": t_adj, "ratio": t_adj / t_fwd, "adjoint_first_call_s": t_first,
             "loss_per_member": [float(x) for x in loss]}
    return {"forward": part1, "batched_adjoint": part2}


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus %d" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    from dgrain import configs, native

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    wl = configs.coastal(N=a.N, start="flood")
    s = native.Solver(wl.mesh, wl.solver_params())
    s.set_state(wl.zeta, wl.qx, wl.qy)
    s.run(a.warmup)
    s.sync()
    # pick the fastest bit-identical stage-kernel variant for this mesh
    # (dgr_autotune restores the state it started from)
    if a.variant < 0:
        variant, tune_ms = s.autotune(10)
    else:
        s.set_variant(a.variant)
        variant, tune_ms = a.variant, None
    # and the workgroup size of the element and node kernels
    if a.block <= 0:
        block, block_ms = s.autotune_block(10)
    else:
        s.set_block(a.block)
        block, block_ms = a.block, None
    # the headline: HIP events on the handle's stream around the timed steps
    # (no per-kernel events inside), wall clock beside them
    rep, wall, ev = timed_run(s, a.steps)
    ne = wl.num_el
    dof_updates = 9.0 * ne * 2 * a.steps
    value = dof_updates / ev
    stage_s = ev / (2 * a.steps)                 # one RK stage, all its kernels
    stage_gbs = STAGE_BYTES_PER_EL * ne / stage_s / 1e9
    # the flux/RHS/update pair alone, in a separate profiled pass (per-kernel
    # event pairs slow a run down, so they stay out of the headline)
    s.profile_begin()
    s.run(min(a.steps, 50))
    k_ms, k_n = s.profile_end()
    k_avg_s = (k_ms / k_n) * 1e-3
    pair_gbs = KSTAGE_BYTES_PER_EL * ne / k_avg_s / 1e9
    traffic, valu = load_pair(variant) if a.N == 512 else (None, None)
    cnt = load_counters(variant) if a.N == 512 else None
    stage_roof = {"kernel": "one RK stage (%s)" % STAGE_ALL[variant],
                  "achieved": stage_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": stage_gbs / HBM_PEAK_GBS,
                  "traffic": None, "avg_launch_ms": stage_s * 1e3, "launches": 2 * a.steps,
                  "bytes_per_element": STAGE_BYTES_PER_EL, "bound": "hbm"}
    fp64 = None
    if cnt is not None:
        ps = cnt["per_stage"]
        hbm_frac = ps["bytes"] / stage_s / 1e9 / HBM_PEAK_GBS
        valu_frac = ps["valu_insts"] / stage_s / 1e9 / VALU_PEAK_GWIS
        flops = ps["f64_flops"] / stage_s      # lane FLOPs: 64 x the per-wave counters
        stage_roof.update({"traffic": ps["bytes"], "traffic_frac_of_peak": hbm_frac, "valu_issue_frac": valu_frac,
                           "bound": "valu" if valu_frac > hbm_frac else "hbm"})
        fp64 = {"achieved": flops / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": flops / 1e12 / FP64_PEAK_TFLOPS, "flops_per_stage": ps["f64_flops"],
                "flops_per_element_stage": ps["f64_flops"] / ne,
                "counters": "64 x (SQ_INSTS_VALU_FLOPS_FP64 + SQ_INSTS_VALU_FLOPS_FP64_TRANS): the counters count per wave-instruction"}
    out = {
        "metric": "FP64 DOF-updates/s per RK stage",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ev / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "coastal_inundation_flood_start_%dx%d" % (a.N, a.N), "triangles": ne,
                   "dofs": 9 * ne, "p": 1, "rk": "SSP-RK2", "cfl": 0.3, "wetdry": True, "limiter": True,
                   "tide_m": 1.5, "sim_time_s": rep.t},
        "roofline": stage_roof,
        "fp64": fp64,
        "wall_ms_per_step": wall / a.steps * 1e3,
        # the flux/RHS/update kernel pair alone (HIP events around each
        # launch of the pair, profiled pass)
        "roofline_pair": {"bound": "valu", "kernel": STAGE_KERNELS[variant], "achieved": pair_gbs,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": pair_gbs / HBM_PEAK_GBS,
                          "traffic": traffic, "avg_launch_ms": k_avg_s * 1e3, "launches": k_n,
                          "bytes_per_element": KSTAGE_BYTES_PER_EL},
        # FP64 VALU issue of the pair (wave-instructions/s against 1024 SIMDs x
        # 2.4 GHz / 4 cycles), instruction count from the PMC pass
        "valu_issue": None if valu is None else {
            "achieved": valu / k_avg_s / 1e9, "peak": VALU_PEAK_GWIS, "unit": "G wave-instructions/s",
            "frac": valu / k_avg_s / 1e9 / VALU_PEAK_GWIS, "insts_per_launch": valu},
        "stage_variant": {"chosen": variant, "autotune_ms_per_step": tune_ms, "block": block,
                          "block_autotune_ms_per_step_64_128_256": block_ms},
        "ensemble_speedup": None,
        "adjoint_forward_cost": None,
    }
    el = ev
    s.close()
    progress("single instance: %.3e DOF-updates/s, %.4f ms/step" % (value, el / a.steps * 1e3))
    if not a.no_wetdry:
        out["wetdry"] = wetdry_bench(a, variant)
        progress("wet/dry leg: %.3e DOF-updates/s, %.4f ms/step" % (out["wetdry"]["value"],
                                                                   out["wetdry"]["ms_per_step"]))
    if not a.no_ensemble:
        e = ensemble_bench(a, variant)
        out["ensemble"] = e
        out["ensemble_speedup"] = e["speedup"]
        progress("ensemble: speedup %.2f" % e["speedup"])
    if not a.no_adjoint:
        ad = adjoint_bench(a, variant)
        out["adjoint"] = ad
        out["adjoint_forward_cost"] = ad["ratio"]
        progress("adjoint: %.2f x forward" % ad["ratio"])
    if not a.no_large:
        out["large"] = large_bench(a, variant)
        progress("large: %.3e DOF-updates/s; batched adjoint %.2f x forward"
                 % (out["large"]["forward"]["value"], out["large"]["batched_adjoint"]["ratio"]))
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_seconds)
        # the host's share of cores per GPU is 16 on the measurement machines
        nt = min(16, os.cpu_count() or 1)
        if nt > 1:
            out["cpu_baseline_threads"] = cpu_baseline(wl, a.cpu_seconds / 2, threads=nt)
        out["cpu_baseline_reference_flags"] = cpu_baseline(wl, a.cpu_seconds, ref_flags=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
