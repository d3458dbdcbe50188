// dgr_handle.h -- the handle of libdgrain and the host-side pieces its C ABI
// files share (dgr_api.hip: forward; dgr_adjoint_api.hip: gradients).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../../include/dgrain.h"
#include "dgr_device.h"

namespace dgr {
int launch_stage(int variant, int mode, bool wetdry, const MeshDev &m, const PhysDev &ph, const double *win,
                 double *wout, const double *wbase, const double *bzeta, const double *bqn, const int *wd,
                 double *means, double *fhat, Ctrl *ctrl, int slot, double *dbg_rhs, double *dbg_fhat,
                 hipStream_t s, bool ties = false);
hipError_t launch_stage_persist(int mode, bool wetdry, const MeshDev &m, const PhysDev &ph, const PersistArgs &a,
                                hipStream_t s);
void launch_means(const MeshDev &m, const double *w, double *means, hipStream_t s);
void launch_edge_geometry(const MeshDev &m, double4 *egeo, hipStream_t s);
void launch_node_minmax(const MeshDev &m, const double *means, double *nodemm, int *nodearg, hipStream_t s);
void launch_limit(bool pdop, const MeshDev &m, const PhysDev &ph, const double *win, double *wout,
                  const double *nodemm, unsigned char *modify, Ctrl *ctrl, int stage, const int *wd_in, int *wd_out,
                  hipStream_t s, int step_adj = 0);
// stage-2 node bounds with the step end of every member (k_step_end, next = 1)
// in the same launch
void launch_node_minmax_end(const MeshDev &m, const PhysDev &ph, const double *means, double *nodemm, int *nodearg,
                            Ctrl *c, int kind, StepRec *rec_end, StepRec *rec_next, bool ties, hipStream_t s);
void launch_pdop(const MeshDev &m, const PhysDev &ph, double *w, Ctrl *ctrl, hipStream_t s);
void launch_wd_status(const MeshDev &m, const PhysDev &ph, const double *w, const int *wd_in, int *wd_out,
                      hipStream_t s);
void launch_step_begin(const MeshDev &m, Ctrl *c, hipStream_t s);
void launch_couple_dgshed(const MeshDev &m, const PhysDev &ph, const double *w, double *bzeta, double *bqn, int n,
                          int *spec, double *vals, hipStream_t s);
void launch_couple_ends(const MeshDev &m, const double *w, double *bzeta, double *bqn, int nends, const int *spec,
                        const double *vals, double *ghost, hipStream_t s);
void launch_step_end(const MeshDev &m, const PhysDev &ph, Ctrl *c, int kind, bool next, hipStream_t s,
                     StepRec *rec_end = nullptr, StepRec *rec_next = nullptr, bool ties = false);
void launch_lam_reduce(const MeshDev &m, const PhysDev &ph, Ctrl *c, int slot, int kind, hipStream_t s);
void launch_step_mark(const MeshDev &m, Ctrl *c, double dt, hipStream_t s);
void launch_reset_maxlam(const MeshDev &m, Ctrl *c, hipStream_t s);
void launch_math_selftest(int n, const double *x, const double *y, double *out, hipStream_t s);

// adjoint sweep (dgr_adjoint_kernels.h)
void launch_adj_limit(const MeshDev &m, const PhysDev &ph, const double *U, const double *nodemm, const double *Sb,
                      const double *slotb, const double *zbe, double *zvb, double *Ub, double *vmb, AdjCtrl *actl,
                      const AdjScalarArgs *sa, hipStream_t s);
void launch_adj_node_sums(const MeshDev &m, const double *vmb, double *nsum, hipStream_t s);
void launch_adj_rhs(int mode, const MeshDev &m, const PhysDev &ph, const double *W, const double *B,
                    const double *U, const double *Ubown, const double *nsum, const int *nodearg, const int *wd,
                    const StepRec *rec, double *Inb, double *Bb, double *Fhb, double *zvb, double *nvb,
                    double *dtparts, hipStream_t s);
void launch_count_notwet(const MeshDev &m, const int *wd, int *out, hipStream_t s);
void launch_adj_edge(const MeshDev &m, const PhysDev &ph, const double *win, const double *bzeta, const double *bqn,
                     const int *wd, const double *Fhb, const StepRec *rec, const StepRec *rec_lam, AdjCtrl *actl,
                     double *slotb, double *zbe, double *tideparts, int with_lambda, const AdjEdgeQueue &eq,
                     hipStream_t s);
void launch_adj_ties(const MeshDev &m, const PhysDev &ph, const double *win, const double *bzeta, const double *bqn,
                     const int *wd, const StepRec *rec, const StepRec *rec_lam, AdjCtrl *actl, hipStream_t s);
void launch_adj_combine(const MeshDev &m, double *Dst, const double *slotb, const double *zbe, double *zvb,
                        hipStream_t s);
void launch_adj_scalar(const MeshDev &m, const AdjScalarArgs &sa, AdjCtrl *actl, hipStream_t s);
void launch_adj_gauges(const MeshDev &m, int ngauge, const int *gauge_el, const double *hval, const double *obs,
                       double *Wb, double *zvb, hipStream_t s);
// member mid's gauge g goes to hval[mid * mstride + g]; mstride 0: ngauge
void launch_gauge_sample(const MeshDev &m, int ngauge, const int *gauge_el, const double *w, double *hval,
                         const double *obs, AdjCtrl *actl, hipStream_t s, int mstride = 0);
void launch_step_record(const MeshDev &m, const Ctrl *ctrls, StepRec *rec, int at_end, hipStream_t s);
void launch_adj_gather(const MeshDev &m, const double *zvb, const double *nvb, double *gz, double *gn, hipStream_t s);
}  // namespace dgr

using namespace dgr;

struct MemberHost {
    int ne, nn, ned;              // the member's own sizes
    int el_off, nd_off, ed_off;   // offsets into the concatenated arrays
    dgr_params prm;
    double min_edge;
    std::vector<double> hyd[3];   // hydrograph of code c+1: t, zeta_b, qn_b (3n values)
    double infil_K = 0.0, infil_P = 0.0;   // Green-Ampt K and psi * delta_theta (K = 0: off)
};

struct dgr_handle {
    int nmem;
    int pitch, nn, ned_pad;       // padded element count, total nodes, padded edge count
    int device;
    hipStream_t stream;
    std::vector<MemberHost> mem;
    PhysDev ph;
    MeshDev m;
    int record_modify;
    // device buffers
    double4 *d_nodes;
    double *d_zn;   // per-node bed, a copy of d_nodes[i].z
    int *d_ev, *d_nb, *d_info, *d_eid, *d_n2e_off, *d_n2e;
    int *d_nl_off, *d_nl, *d_loc_off;   // node groups' element lists, entries' positions in them
    unsigned *d_locw;
    int4 *d_edges;   // {el_in, el_ex, va, vb} per edge
    int *d_einfo;    // packed edge facts (ED_*/EI_* bits)
    double4 *d_egeo; // canonical edge geometry (MeshDev::egeo), refreshed when the nodes change
    MemberDev *d_mem;
    int *d_eblk_mem, *d_dblk_mem;
    double *d_bzeta, *d_bqn;
    double *d_w, *d_w1, *d_means, *d_nodemm;
    double *d_dbg;   // 18 planes: rhs (9) then fhat (9)
    double *d_fhat;  // 9 planes of edge fluxes (split kernel variants)
    int *d_wd;
    unsigned char *d_modify;   // [2][pitch]
    Ctrl *d_ctrl;              // [nmem]
    unsigned long long *d_lamparts;   // [2][maxparts] per-workgroup max |lambda|
    int *d_nanparts;                  // [2][maxparts] last NaN operand per edge workgroup
    unsigned long long *d_tiemax;     // [maxparts] tie partials of a taped stage 2
    int *d_tiecnt;
    bool ties_last;                   // the last enqueued step counted max-|lambda| ties
    int *d_wdn;      // [pitch] wet/dry flags for the next step's start (dgr_run)
    double *d_hydro; // hydrograph tables of all members (MeshDev::hydro), or null
    int variant;     // 0 fused, 1 split element-centric, 2 split edge-centric
    int kind[2];     // partials space of the last flux launch of each stage
    // host staging
    std::vector<double> h_planes;
    // optional HIP-event timing of the flux/RHS/update launches
    int profiling;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used;
    // one captured step (dgr_set_graph)
    int use_graph;
    int graph_variant;
    int graph_wg;    // MeshDev::wg the graph was captured with
    hipGraphExec_t graph_exec;
    // HBM the adjoint may keep taped segments in (dgr_set_tape_budget):
    // < 0 the free device memory less 1/16 of it, 0 recompute every segment
    long long tape_budget = -1;
    // the adjoint's tape space, kept for the next call (a gradient loop
    // calls with the same sizes)
    char *d_tape = nullptr;
    size_t tape_bytes = 0;
    // channel-coupling staging (dgr_couple_channel_ends), grown on demand
    int *d_cpl_spec = nullptr;
    double *d_cpl_vals = nullptr;
    size_t cpl_spec_n = 0, cpl_vals_n = 0;
};

#define DGR_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "libdgrain: %s failed: %s\n", #call, hipGetErrorString(e_)); \
            return DGR_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)


template <typename T>
static inline hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
}

// Buffers of one step.  The plain step works in place (w0 = u2 = wout =
// d_w, u1 = p1 = d_w1, every flag set in d_wd); a taped step keeps every
// stage output and flag set in its own buffer for the reverse sweep.
struct StepBufs {
    double *w0;     // state at the step's start
    double *u1;     // stage 1 output before the limiter
    double *p1;     // stage 1 output after limiter/PDop (stage 2 input)
    double *u2;     // stage 2 output before the limiter
    double *wout;   // stage 2 output after limiter/PDop (next step's start)
    const int *wdprev;   // wet/dry flags on entry (the previous step's stage-2 flags)
    int *wda;       // wet/dry flags stage 1 uses
    int *wdb;       // wet/dry flags stage 2 uses
    int *wdnext;    // if set: the stage-2 limiter also writes the next step's
                    // start flags (wetDryStatus2D of wout from wdb) here
    bool wda_ready; // wda already holds this step's start flags (a previous
                    // step wrote them through wdnext): no status launch
    // if set: each stage's node bounds [6][nn] (and the elements attaining
    // them) go here instead of the handle's scratch buffer
    double *nodemm1, *nodemm2;
    int *nodearg1, *nodearg2;
    // a taped step: the stage-2 edge kernel also counts the max-|lambda|
    // ties the reverse needs
    bool taped;
};

// fuse_end: the step end (k_step_end, next = 1, with the optional step
// records) rides on the stage-2 node bounds launch; the caller then launches
// no step end
void enqueue_stages(dgr_handle *h, const StepBufs *b, bool fuse_end = false, StepRec *rec_end = nullptr,
                    StepRec *rec_next = nullptr);
void enqueue_step(dgr_handle *h, const StepBufs *b);
// the steps of a run: the prologue once (step bookkeeping, start flags into
// d_wdn), then run steps that chain the start flags through the stage-2
// limiter and end with the next step's bookkeeping.  b = nullptr: the
// handle's own buffers (wda = d_wdn, wdb = d_wd, in place).
void enqueue_run_prologue(dgr_handle *h);
void enqueue_run_step(dgr_handle *h, const StepBufs *b = nullptr, StepRec *rec_end = nullptr,
                      StepRec *rec_next = nullptr);
dgr_status read_ctrl(dgr_handle *h, int mi, Ctrl *c);
dgr_status any_error(dgr_handle *h);
unsigned char *modify_ptr(dgr_handle *h, int stage);
