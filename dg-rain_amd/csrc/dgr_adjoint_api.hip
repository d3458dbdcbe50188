// dgr_adjoint_api.hip -- C ABI of the gradient: a gauge-misfit loss over a
// forward run and its adjoint with respect to nodal depth z and Manning n
// (per member), by checkpointing and recomputing the forward.
//
// Forward: nsteps steps from the handle's state, a checkpoint (state, wet/dry
// flags, control block) every `checkpoint_every` steps, the gauge depths
// sampled after every `sample_every`-th step.  Reverse: for each segment from
// the last, restore its checkpoint, recompute its steps with a tape (each
// step's start state, the two stage outputs before and after the limiter and
// the wet/dry flags), then run the reverse stages step by step back to the
// segment start.  The same kernels recompute the same bits, so the reverse
// sees the forward pass's active sets exactly.  The handle's state is
// restored to the start state on return.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "dgr_handle.h"

namespace {

struct AdjBuffers {
    std::vector<void *> owned;
    template <typename T>
    hipError_t alloc(T **p, size_t n)
    {
        hipError_t e = dalloc(p, n);
        if (e == hipSuccess) owned.push_back((void *)*p);
        return e;
    }
    ~AdjBuffers()
    {
        for (void *p : owned) (void)hipFree(p);
    }
};

}  // namespace

extern "C" dgr_status dgr_gauge_run(dgr_handle *h, int nsteps, int sample_every, int ngauges, const int *gauge_el,
                                    double *h_out)
{
    if (!h || nsteps < 0 || sample_every <= 0 || ngauges <= 0 || !gauge_el || !h_out) return DGR_ERR_ARG;
    for (int mi = 0; mi < h->nmem; ++mi)
        for (int g = 0; g < ngauges; ++g) {
            const int e = gauge_el[mi * ngauges + g];
            if (e < 0 || e >= h->mem[mi].ne) return DGR_ERR_ARG;
        }
    const int nsamp = nsteps / sample_every;
    const size_t per = (size_t)h->nmem * ngauges;
    AdjBuffers B;
    int *d_gel = nullptr;
    double *d_h = nullptr;
    DGR_HIP(B.alloc(&d_gel, per));
    DGR_HIP(B.alloc(&d_h, per * (nsamp > 0 ? nsamp : 1)));
    DGR_HIP(hipMemcpyAsync(d_gel, gauge_el, per * sizeof(int), hipMemcpyHostToDevice, h->stream));
    if (nsteps > 0) enqueue_run_prologue(h);
    for (int n = 0; n < nsteps; ++n) {
        enqueue_run_step(h);
        if ((n + 1) % sample_every == 0) {
            const int k = (n + 1) / sample_every - 1;
            launch_gauge_sample(h->m, ngauges, d_gel, h->d_w, d_h + (size_t)k * per, nullptr, nullptr, h->stream);
        }
    }
    DGR_HIP(hipGetLastError());
    std::vector<double> hv(per * (nsamp > 0 ? nsamp : 1));
    DGR_HIP(hipMemcpyAsync(hv.data(), d_h, hv.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    // [member][sample][gauge]
    for (int mi = 0; mi < h->nmem; ++mi)
        for (int k = 0; k < nsamp; ++k)
            for (int g = 0; g < ngauges; ++g)
                h_out[((size_t)mi * nsamp + k) * ngauges + g] = hv[(size_t)k * per + (size_t)mi * ngauges + g];
    return any_error(h);
}

// ---------------------------------------------------------------------------
// dgr_gauge_run_noalloc: dgr_gauge_run without device allocation.  Its
// staging (the gauge element ids, then the sampled depths already in the
// output's [member][sample][gauge] order) lives in `scratch`, a fixed array
// in each device's memory for the life of the process.  On the host the
// array's name is not a device address: each device's copy is found once
// with hipGetSymbolAddress (that device current) and only that address is
// handed to copies and kernels.  Every call on a device shares its copy, so
// calls that overlap on one device race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[DGR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kMaxDevices];   // this device's copy of `scratch`

dgr_status scratch_of(int device, unsigned char **out)
{
    if (device < 0 || device >= kMaxDevices) return DGR_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        int cur = -1;
        DGR_HIP(hipGetDevice(&cur));
        if (cur != device) DGR_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            fprintf(stderr, "libdgrain: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return DGR_ERR_HIP;
        }
        g_scratch_dev[device] = (unsigned char *)p;
    }
    *out = g_scratch_dev[device];
    return DGR_OK;
}
}  // namespace

extern "C" dgr_status dgr_gauge_run_noalloc(dgr_handle *h, int nsteps, int sample_every, int ngauges,
                                            const int *gauge_el, double *h_out)
{
    if (!h || nsteps < 0 || sample_every <= 0 || ngauges <= 0 || !gauge_el || !h_out) return DGR_ERR_ARG;
    for (int mi = 0; mi < h->nmem; ++mi)
        for (int g = 0; g < ngauges; ++g) {
            const int e = gauge_el[mi * ngauges + g];
            if (e < 0 || e >= h->mem[mi].ne) return DGR_ERR_ARG;
        }
    const size_t nsamp = (size_t)(nsteps / sample_every);
    const size_t per = (size_t)h->nmem * ngauges;
    const size_t id_bytes = (per * sizeof(int) + 255) / 256 * 256;
    if (id_bytes + per * nsamp * sizeof(double) > DGR_NOALLOC_SCRATCH_BYTES) {
        fprintf(stderr, "libdgrain: dgr_gauge_run_noalloc: %zu gauge ids and %zu samples exceed its %u-byte workspace\n",
                per, per * nsamp, (unsigned)DGR_NOALLOC_SCRATCH_BYTES);
        return DGR_ERR_TOO_LARGE;
    }
    unsigned char *base = nullptr;
    const dgr_status st = scratch_of(h->device, &base);
    if (st != DGR_OK) return st;
    int *d_gel = (int *)base;
    double *d_h = (double *)(base + id_bytes);
    DGR_HIP(hipMemcpyAsync(d_gel, gauge_el, per * sizeof(int), hipMemcpyHostToDevice, h->stream));
    if (nsteps > 0) enqueue_run_prologue(h);
    const int mstride = (int)nsamp * ngauges;
    for (int n = 0; n < nsteps; ++n) {
        enqueue_run_step(h);
        if ((n + 1) % sample_every == 0) {
            const int k = (n + 1) / sample_every - 1;
            launch_gauge_sample(h->m, ngauges, d_gel, h->d_w, d_h + (size_t)k * ngauges, nullptr, nullptr, h->stream,
                                mstride);
        }
    }
    DGR_HIP(hipGetLastError());
    if (per * nsamp > 0)
        DGR_HIP(hipMemcpyAsync(h_out, d_h, per * nsamp * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    // the members' error keys, a few control blocks at a time on the stack
    Ctrl c[8];
    for (int m0 = 0; m0 < h->nmem; m0 += 8) {
        const int nc = std::min(8, h->nmem - m0);
        DGR_HIP(hipMemcpyAsync(c, h->d_ctrl + m0, nc * sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
        DGR_HIP(hipStreamSynchronize(h->stream));
        for (int i = 0; i < nc; ++i)
            if (err_code_of(c[i].err_key)) return DGR_ERR_NAN;
    }
    return DGR_OK;
}

extern "C" dgr_status dgr_set_nodal(dgr_handle *h, int member, const double *z, const double *n)
{
    if (!h || member < 0 || member >= h->nmem || (!z && !n)) return DGR_ERR_ARG;
    const MemberHost &M = h->mem[member];
    std::vector<double4> nodes(M.nn);
    DGR_HIP(hipMemcpyAsync(nodes.data(), h->d_nodes + M.nd_off, M.nn * sizeof(double4), hipMemcpyDeviceToHost,
                           h->stream));
    DGR_HIP(hipStreamSynchronize(h->stream));
    for (int i = 0; i < M.nn; ++i) {
        if (z) nodes[i].z = z[i];
        if (n) nodes[i].w = n[i];
    }
    DGR_HIP(hipMemcpyAsync(h->d_nodes + M.nd_off, nodes.data(), M.nn * sizeof(double4), hipMemcpyHostToDevice,
                           h->stream));
    if (z) DGR_HIP(hipMemcpyAsync(h->d_zn + M.nd_off, z, M.nn * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (z) launch_edge_geometry(h->m, h->d_egeo, h->stream);   // z_edge follows the bed
    DGR_HIP(hipGetLastError());
    DGR_HIP(hipStreamSynchronize(h->stream));
    return DGR_OK;
}

extern "C" dgr_status dgr_set_tape_budget(dgr_handle *h, long long bytes)
{
    if (!h) return DGR_ERR_ARG;
    h->tape_budget = bytes < 0 ? -1 : bytes;
    if (h->d_tape) {
        DGR_HIP(hipStreamSynchronize(h->stream));
        DGR_HIP(hipFree(h->d_tape));
        h->d_tape = nullptr;
        h->tape_bytes = 0;
    }
    return DGR_OK;
}

extern "C" dgr_status dgr_adjoint_gauges(dgr_handle *h, const dgr_adjoint_spec *sp, double *loss, double *grad_z,
                                         double *grad_n)
{
    if (!h || !sp || !loss || sp->nsteps <= 0 || sp->sample_every <= 0 || sp->checkpoint_every <= 0 ||
        sp->ngauges <= 0 || !sp->gauge_el || !sp->obs)
        return DGR_ERR_ARG;
    const int N = sp->nsteps, K = sp->checkpoint_every, ng = sp->ngauges, nm = h->nmem;
    for (int mi = 0; mi < nm; ++mi)
        for (int g = 0; g < ng; ++g) {
            const int e = sp->gauge_el[mi * ng + g];
            if (e < 0 || e >= h->mem[mi].ne) return DGR_ERR_ARG;
        }
    if (h->profiling) return DGR_ERR_ARG;
    // the reverse has no adjoint of a hydrograph's time dependence, nor of
    // the infiltration rate's dependence on t and dt
    for (int mi = 0; mi < h->nmem; ++mi) {
        for (int c = 0; c < 3; ++c)
            if (!h->mem[mi].hyd[c].empty()) return DGR_ERR_UNSUPPORTED;
        if (h->mem[mi].infil_K > 0.0) return DGR_ERR_UNSUPPORTED;
    }
    const int nseg = (N + K - 1) / K, nsamp = N / sp->sample_every;
    const size_t P = h->pitch, W9 = 9 * P, nn = h->nn, per = (size_t)nm * ng;
    const int maxparts = h->ph.maxparts;
    hipStream_t s = h->stream;
    AdjBuffers B;
    // checkpoints, step records, gauges
    double *ckW;
    int *ckWD;
    Ctrl *ckC;
    StepRec *rec;
    int *d_gel;
    double *d_obs, *d_hv;
    AdjCtrl *actl;
    DGR_HIP(B.alloc(&ckW, W9 * nseg));
    DGR_HIP(B.alloc(&ckWD, P * nseg));
    DGR_HIP(B.alloc(&ckC, (size_t)nm * nseg));
    DGR_HIP(B.alloc(&rec, (size_t)nm * N));
    DGR_HIP(B.alloc(&d_gel, per));
    DGR_HIP(B.alloc(&d_obs, per * (nsamp > 0 ? nsamp : 1)));
    DGR_HIP(B.alloc(&d_hv, per * (nsamp > 0 ? nsamp : 1)));
    DGR_HIP(B.alloc(&actl, nm));
    // reverse buffers: slot/zbe pairs per stage (the edge part of a stage's
    // input adjoint is consumed by the limiter reverse that comes next)
    double *wb, *wb0, *pb, *ub, *fhb, *slot1, *slot2, *zbe1, *zbe2, *vmb, *nsum, *zvb, *nvb, *dtp, *tdp, *gz, *gn;
    DGR_HIP(B.alloc(&wb, W9));
    DGR_HIP(B.alloc(&wb0, W9));
    DGR_HIP(B.alloc(&pb, W9));
    DGR_HIP(B.alloc(&ub, W9));
    DGR_HIP(B.alloc(&fhb, W9));
    DGR_HIP(B.alloc(&slot1, W9));
    DGR_HIP(B.alloc(&slot2, W9));
    DGR_HIP(B.alloc(&zbe1, 3 * P));
    DGR_HIP(B.alloc(&zbe2, 3 * P));
    DGR_HIP(B.alloc(&vmb, 18 * P));
    DGR_HIP(B.alloc(&nsum, 8 * nn));
    DGR_HIP(B.alloc(&zvb, 3 * P));
    DGR_HIP(B.alloc(&nvb, 3 * P));
    DGR_HIP(B.alloc(&dtp, 2 * (size_t)maxparts));
    DGR_HIP(B.alloc(&tdp, 2 * (size_t)maxparts));
    DGR_HIP(B.alloc(&gz, nn));
    DGR_HIP(B.alloc(&gn, nn));
    AdjEdgeQueue eq{};
    const bool split = h->ph.wetdry && h->m.ned_pad >= ADJ_EDGE_SPLIT;
    const size_t neq = split ? (size_t)h->m.ned_pad : 256;
    DGR_HIP(B.alloc(&eq.blist, neq));
    DGR_HIP(B.alloc(&eq.tb, neq));
    DGR_HIP(B.alloc(&eq.queue, neq / 256));
    DGR_HIP(B.alloc(&eq.qcount, 2));
    // one stage's node bounds and the elements attaining them, when the
    // tapes leave them out (compact tapes, below)
    double *cnm;
    int *cna;
    DGR_HIP(B.alloc(&cnm, 6 * nn));
    DGR_HIP(B.alloc(&cna, 8 * nn));
    // Tapes: each step's start state, both stage outputs before and after
    // the limiter, both stages' flags, node bounds and the elements attaining
    // them.  The last R segments are taped by the forward pass itself and
    // stay resident (R from the tape budget); the others are recomputed from
    // their checkpoints into slot 0 when the reverse reaches them.  Both
    // ways run the same kernels on the same inputs: the tapes are identical.
    const size_t Wb = W9 * sizeof(double), Ib = P * sizeof(int), Cb = (size_t)nm * sizeof(Ctrl);
    const bool wdon = h->ph.wetdry != 0;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_w = up(Wb * K), b_i = up(Ib * K), b_nm = up(12 * nn * K * sizeof(double)),
                 b_na = up(16 * nn * K * sizeof(int));
    const size_t seg_full = 4 * b_w + 2 * b_i + b_nm + b_na;
    // compact tapes: states and flags only; the reverse recomputes each
    // stage's node bounds from the taped pre-limiter state (means, then the
    // node kernel: the forward's kernels on the forward's inputs, so the
    // same bits).  Chosen when they keep more segments resident than full
    // tapes do.
    const size_t seg_compact = 4 * b_w + 2 * b_i;
    long long budget = h->tape_budget;
    if (budget < 0) {
        // the free memory (every other buffer of the call is allocated by
        // now), counting the tape space the handle holds, less a reserve of
        // 1/16 of the device memory
        size_t fr = 0, tot = 0;
        budget = hipMemGetInfo(&fr, &tot) == hipSuccess
                     ? std::max<long long>(0, (long long)(fr + h->tape_bytes) - (long long)(tot / 16))
                     : 0;
    }
    const int R_full = (int)std::min<long long>(nseg, budget / (long long)seg_full);
    const int R_compact = (int)std::min<long long>(nseg, budget / (long long)seg_compact);
    const bool compact = R_compact > R_full;
    const size_t seg_bytes = compact ? seg_compact : seg_full;
    int R = compact ? R_compact : R_full;
    if (h->tape_bytes < seg_bytes * (size_t)(R > 1 ? R : 1)) {
        // grow the handle's tape space (kept for the next call)
        if (h->d_tape) {
            DGR_HIP(hipStreamSynchronize(s));
            DGR_HIP(hipFree(h->d_tape));
            h->d_tape = nullptr;
            h->tape_bytes = 0;
        }
        for (;;) {
            const size_t want = seg_bytes * (size_t)(R > 1 ? R : 1);
            if (hipMalloc((void **)&h->d_tape, want) == hipSuccess) {
                h->tape_bytes = want;
                break;
            }
            (void)hipGetLastError();
            h->d_tape = nullptr;
            if (R <= 1) return DGR_ERR_HIP;
            R /= 2;
        }
    }
    char *const tape = h->d_tape;
    struct Tape {
        double *w0, *u1, *p1, *u2, *nm;
        int *wa, *wb, *na;
    };
    auto slot = [&](int j) {
        char *b = tape + seg_bytes * (size_t)j;
        Tape t;
        t.w0 = (double *)b; b += b_w;
        t.u1 = (double *)b; b += b_w;
        t.p1 = (double *)b; b += b_w;
        t.u2 = (double *)b; b += b_w;
        t.wa = (int *)b; b += b_i;
        t.wb = (int *)b; b += b_i;
        t.nm = compact ? nullptr : (double *)b; b += compact ? 0 : b_nm;
        t.na = compact ? nullptr : (int *)b;
        return t;
    };
    const int first_res = nseg - R;   // segments >= first_res stay taped from the forward
    // step n's buffers in tape T (k = n - segment start; last: the segment's
    // last step, whose output is the handle's state)
    auto taped_step = [&](const Tape &T, int k, bool last) {
        StepBufs bf{};
        bf.w0 = T.w0 + W9 * k;
        bf.u1 = T.u1 + W9 * k;
        bf.p1 = T.p1 + W9 * k;
        bf.u2 = T.u2 + W9 * k;
        bf.wout = last ? h->d_w : T.w0 + W9 * (k + 1);
        if (wdon) {
            bf.wdprev = k == 0 ? h->d_wd : T.wb + P * (k - 1);
            bf.wda = T.wa + P * k;
            bf.wdb = T.wb + P * k;
            bf.wdnext = last ? h->d_wdn : T.wa + P * (k + 1);
            bf.wda_ready = true;
        } else {
            bf.wdprev = bf.wda = bf.wdb = h->d_wd;   // never written without wet/dry
        }
        if (!compact) {
            bf.nodemm1 = T.nm + 6 * nn * (2 * k);
            bf.nodemm2 = T.nm + 6 * nn * (2 * k + 1);
            bf.nodearg1 = T.na + 8 * nn * (2 * k);
            bf.nodearg2 = T.na + 8 * nn * (2 * k + 1);
        }
        bf.taped = true;
        return bf;
    };
    DGR_HIP(hipMemcpyAsync(d_gel, sp->gauge_el, per * sizeof(int), hipMemcpyHostToDevice, s));
    {
        // obs arrive [member][sample][gauge]; device keeps [sample][member][gauge]
        std::vector<double> o(per * (nsamp > 0 ? nsamp : 1));
        for (int mi = 0; mi < nm; ++mi)
            for (int k = 0; k < nsamp; ++k)
                for (int g = 0; g < ng; ++g)
                    o[(size_t)k * per + (size_t)mi * ng + g] = sp->obs[((size_t)mi * nsamp + k) * ng + g];
        DGR_HIP(hipMemcpyAsync(d_obs, o.data(), o.size() * sizeof(double), hipMemcpyHostToDevice, s));
        DGR_HIP(hipMemsetAsync(actl, 0, nm * sizeof(AdjCtrl), s));
        DGR_HIP(hipStreamSynchronize(s));
    }

    // forward: checkpoints of the recomputed segments (and of the start, to
    // restore the handle), taped steps in the resident ones; each step's end
    // records its max |lambda| and the next step's t, dt and tide
    enqueue_run_prologue(h);
    launch_step_record(h->m, h->d_ctrl, rec, 0, s);
    for (int n = 0; n < N; ++n) {
        const int c = n / K, k = n % K;
        const int n1 = (c + 1) * K < N ? (c + 1) * K : N;
        StepRec *re = rec + (size_t)nm * n, *rx = n + 1 < N ? rec + (size_t)nm * (n + 1) : nullptr;
        if (k == 0 && (c < first_res || c == 0)) {
            DGR_HIP(hipMemcpyAsync(ckW + W9 * c, h->d_w, Wb, hipMemcpyDeviceToDevice, s));
            DGR_HIP(hipMemcpyAsync(ckWD + P * c, h->d_wd, Ib, hipMemcpyDeviceToDevice, s));
            DGR_HIP(hipMemcpyAsync(ckC + (size_t)nm * c, h->d_ctrl, Cb, hipMemcpyDeviceToDevice, s));
        }
        const double *wout = h->d_w;
        if (c < first_res) {
            enqueue_run_step(h, nullptr, re, rx);
        } else {
            const Tape T = slot(c - first_res);
            if (k == 0) {
                // the segment's start state and stage-1 flags (the previous
                // step's stage-2 limiter or the run prologue wrote them)
                DGR_HIP(hipMemcpyAsync(T.w0, h->d_w, Wb, hipMemcpyDeviceToDevice, s));
                if (wdon) DGR_HIP(hipMemcpyAsync(T.wa, h->d_wdn, Ib, hipMemcpyDeviceToDevice, s));
            }
            const StepBufs bf = taped_step(T, k, n + 1 == n1);
            enqueue_run_step(h, &bf, re, rx);
            // the next step's entry flags, where a plain step leaves them
            if (wdon && n + 1 == n1) DGR_HIP(hipMemcpyAsync(h->d_wd, T.wb + P * k, Ib, hipMemcpyDeviceToDevice, s));
            wout = bf.wout;
        }
        if ((n + 1) % sp->sample_every == 0) {
            const int j = (n + 1) / sp->sample_every - 1;
            launch_gauge_sample(h->m, ng, d_gel, wout, d_hv + (size_t)j * per, d_obs + (size_t)j * per, actl, s);
        }
    }
    DGR_HIP(hipGetLastError());
    dgr_status fst = any_error(h);
    auto restore = [&](int c) -> dgr_status {
        DGR_HIP(hipMemcpyAsync(h->d_w, ckW + W9 * c, Wb, hipMemcpyDeviceToDevice, s));
        DGR_HIP(hipMemcpyAsync(h->d_wd, ckWD + P * c, Ib, hipMemcpyDeviceToDevice, s));
        DGR_HIP(hipMemcpyAsync(h->d_ctrl, ckC + (size_t)nm * c, Cb, hipMemcpyDeviceToDevice, s));
        return DGR_OK;
    };
    if (fst != DGR_OK) {
        restore(0);
        DGR_HIP(hipStreamSynchronize(s));
        return fst;
    }

    // reverse.  Taped steps of the edge-kernel variant recorded their
    // max-|lambda| tie counts (h->kind[1] is the variant's partials space)
    const bool rec_ties = h->kind[1] == PARTS_EDGE;
    AdjScalarArgs pending{};
    DGR_HIP(hipMemsetAsync(wb, 0, Wb, s));
    DGR_HIP(hipMemsetAsync(zvb, 0, 3 * P * sizeof(double), s));
    DGR_HIP(hipMemsetAsync(nvb, 0, 3 * P * sizeof(double), s));
    if (split) {
        // the split edge reverse pays off only where nearly every element is
        // wet (its second launch works through the listed edges' workgroups
        // a few hundred at a time): decided on the flags the reverse starts
        // from, at most 1 in 1000 elements not wet
        int notwet = 0;
        size_t nel = 0;
        for (int mi = 0; mi < nm; ++mi) nel += (size_t)h->mem[mi].ne;
        DGR_HIP(hipMemsetAsync(eq.qcount, 0, sizeof(int), s));
        launch_count_notwet(h->m, h->d_wd, eq.qcount, s);
        DGR_HIP(hipGetLastError());
        DGR_HIP(hipMemcpyAsync(&notwet, eq.qcount, sizeof(int), hipMemcpyDeviceToHost, s));
        DGR_HIP(hipStreamSynchronize(s));
        eq.split = (size_t)notwet * 1000 <= nel;
    }
    DGR_HIP(hipMemsetAsync(eq.qcount, 0, 2 * sizeof(int), s));
    bool have_slot1 = false;
    for (int c = nseg - 1; c >= 0; --c) {
        const int n0 = c * K, n1 = (n0 + K < N) ? n0 + K : N;
        const Tape T = slot(c >= first_res ? c - first_res : 0);
        if (c < first_res) {
            // recompute the segment into slot 0: the checkpoint goes to step
            // 0's start, every step reads its slot and writes the next one
            // (no copies); the stage-2 limiter of each step writes the next
            // step's start flags
            DGR_HIP(hipMemcpyAsync(T.w0, ckW + W9 * c, Wb, hipMemcpyDeviceToDevice, s));
            DGR_HIP(hipMemcpyAsync(h->d_wd, ckWD + P * c, Ib, hipMemcpyDeviceToDevice, s));
            DGR_HIP(hipMemcpyAsync(h->d_ctrl, ckC + (size_t)nm * c, Cb, hipMemcpyDeviceToDevice, s));
            launch_step_begin(h->m, h->d_ctrl, s);
            if (wdon) launch_wd_status(h->m, h->ph, T.w0, h->d_wd, T.wa, s);
            // (the step records are written again, with the same values, and
            // with the tie counts of the taped edge kernel)
            for (int n = n0; n < n1; ++n) {
                const StepBufs bf = taped_step(T, n - n0, n + 1 == n1);
                enqueue_run_step(h, &bf, rec + (size_t)nm * n, n + 1 < N ? rec + (size_t)nm * (n + 1) : nullptr);
            }
        }
        for (int n = n1 - 1; n >= n0; --n) {
            const int k = n - n0;
            double *w0 = T.w0 + W9 * k, *u1 = T.u1 + W9 * k, *p1 = T.p1 + W9 * k, *u2 = T.u2 + W9 * k;
            const int *wda = wdon ? T.wa + P * k : h->d_wd, *wdb = wdon ? T.wb + P * k : h->d_wd;
            const StepRec *rn = rec + (size_t)nm * n;
            if ((n + 1) % sp->sample_every == 0) {
                const int j = (n + 1) / sp->sample_every - 1;
                launch_adj_gauges(h->m, ng, d_gel, d_hv + (size_t)j * per, d_obs + (size_t)j * per, wb, zvb, s);
            }
            const double *nm1 = compact ? cnm : T.nm + 6 * nn * (2 * k);
            const double *nm2 = compact ? cnm : T.nm + 6 * nn * (2 * k + 1);
            const int *na1 = compact ? cna : T.na + 8 * nn * (2 * k);
            const int *na2 = compact ? cna : T.na + 8 * nn * (2 * k + 1);
            // a stage's node bounds from its taped pre-limiter state
            auto node_bounds = [&](const double *u) {
                launch_means(h->m, u, h->d_means, s);
                launch_node_minmax(h->m, h->d_means, cnm, cna, s);
            };
            // stage 2: input p1, base w0, output u2 -> limited -> w^{n+1} (adjoint wb + slot1)
            // (the scalar pass of step n+1 rides on this launch)
            if (compact) node_bounds(u2);
            launch_adj_limit(h->m, h->ph, u2, nm2, wb, have_slot1 ? slot1 : nullptr, have_slot1 ? zbe1 : nullptr,
                             zvb, ub, vmb, actl, &pending, s);
            pending.active = 0;
            launch_adj_node_sums(h->m, vmb, nsum, s);
            launch_adj_rhs(2, h->m, h->ph, p1, w0, u2, ub, nsum, na2, wdb, rn, pb, wb0, fhb, zvb, nvb,
                           dtp + maxparts, s);
            // the max-|lambda| adjoint (k_adj_scalar of step n+1) is shared
            // by the edges whose operand attains the recorded max; the taped
            // edge kernel counted them, the element-partials variants count
            // them here
            int lam_mode = 0;
            if (n + 1 < N) {
                lam_mode = rec_ties ? 2 : 1;
                if (!rec_ties) launch_adj_ties(h->m, h->ph, p1, h->d_bzeta, h->d_bqn, wdb, rn, rn, actl, s);
            }
            launch_adj_edge(h->m, h->ph, p1, h->d_bzeta, h->d_bqn, wdb, fhb, rn, rn, actl, slot2, zbe2, tdp + maxparts,
                            lam_mode, eq, s);
            // stage 1: input w0, output u1 -> limited -> p1 (adjoint pb + slot2)
            // (the stage-2 kernels that read cna are ahead on the stream)
            if (compact) node_bounds(u1);
            launch_adj_limit(h->m, h->ph, u1, nm1, pb, slot2, zbe2, zvb, ub, vmb, actl, nullptr, s);
            launch_adj_node_sums(h->m, vmb, nsum, s);
            launch_adj_rhs(1, h->m, h->ph, w0, nullptr, u1, ub, nsum, na1, wda, rn, nullptr, wb0, fhb, zvb, nvb,
                           dtp, s);
            launch_adj_edge(h->m, h->ph, w0, h->d_bzeta, h->d_bqn, wda, fhb, rn, rn, actl, slot1, zbe1, tdp, 0, eq, s);
            have_slot1 = true;
            // this step's scalar pass (dt and tide adjoints to the previous
            // step's max |lambda|): inside the next step's first launch, or
            // a launch of its own after the last step
            pending = AdjScalarArgs{dtp, tdp, rn, n > 0 ? rn - nm : rn, maxparts, n > 0 ? 1 : 0, rec_ties ? 1 : 0, 1};
            if (n == 0) {
                launch_adj_scalar(h->m, pending, actl, s);
                pending.active = 0;
            }
            double *t = wb;
            wb = wb0;
            wb0 = t;
        }
        DGR_HIP(hipGetLastError());
    }
    // the last stage-1 edge part: its z_edge adjoints (the state adjoint of
    // the start is not returned)
    launch_adj_combine(h->m, wb, slot1, zbe1, zvb, s);
    launch_adj_gather(h->m, zvb, nvb, gz, gn, s);
    DGR_HIP(hipGetLastError());
    dgr_status st = restore(0);
    if (st) return st;
    std::vector<AdjCtrl> a(nm);
    DGR_HIP(hipMemcpyAsync(a.data(), actl, nm * sizeof(AdjCtrl), hipMemcpyDeviceToHost, s));
    std::vector<double> hz(nn), hn(nn);
    DGR_HIP(hipMemcpyAsync(hz.data(), gz, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    DGR_HIP(hipMemcpyAsync(hn.data(), gn, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    DGR_HIP(hipStreamSynchronize(s));
    int unsupported = 0;
    size_t o = 0;
    for (int mi = 0; mi < nm; ++mi) {
        loss[mi] = a[mi].loss;
        unsupported |= a[mi].unsupported;
        const MemberHost &M = h->mem[mi];
        for (int i = 0; i < M.nn; ++i) {
            if (grad_z) grad_z[o + i] = hz[M.nd_off + i];
            if (grad_n) grad_n[o + i] = hn[M.nd_off + i];
        }
        o += M.nn;
    }
    return unsupported ? DGR_ERR_UNSUPPORTED : DGR_OK;
}
