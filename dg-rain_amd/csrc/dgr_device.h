// dgr_device.h -- device-side data layout shared by the kernels and the host
// code of libdgrain.
//
// A handle holds one or more independent instances ("members": an ensemble of
// meshes, each with its own forcing, time and dt).  Their elements, nodes and
// edges are concatenated; each member's element range and edge range start on
// a 256-entry boundary, so every workgroup of an element or edge kernel
// belongs to exactly one member and reads that member's parameters and
// control block with scalar loads.  A single instance is the one-member case.
//
// HBM layout (all members):
//   state        double [9][pitch]   planes zeta0..2, qx0..2, qy0..2 (plane f*3+k
//                                     holds the reference's field f at el*3+k)
//   nodes        double4[nn]         {x, y, z, n}: one 32-byte gather per vertex
//   ev           int    [3][pitch]   EltoVert, SoA (global node ids)
//   nb           int    [3][pitch]   neighbour element across local edge k
//                                     (the element itself on a boundary edge)
//   info         int    [3][pitch]   packed edge facts, see EI_* below
//   eid          int    [3][pitch]   global edge id (boundary data lookup)
//   means        double [3][pitch]   element averages of the updated state
//   nodemm       double [6][nn]      node min (zeta,qx,qy) then max
//   n2e_off/n2e  node -> element CSR in ascending element order; an entry is
//                el | k << 30, k the node's local vertex index in el
//   nl_off/nl    per group of 64 nodes (one wave of the node kernels): the
//                distinct elements of the group's CSR entries, ascending
//   loc_off/locw per node: the positions of its CSR entries in its group's
//                list, three 10-bit positions per word, in CSR order
//   edges        int4   [ned_pad]    {el_in, el_ex, va, vb}: EdgtoEls and el_in's
//                                     vertices of the edge in el_in's order
//   einfo        int    [ned_pad]    packed edge facts, see ED_* below
//   egeo         double4[ned_pad]    {nx, ny, z_edge, 1/(nx*ty - ny*tx)} of each
//                                     edge in el_in's orientation, from the nodes
//                                     (one coalesced 32-byte load per edge)
//   eblk_mem     int    [pitch/256]  member of each element workgroup
//   dblk_mem     int    [ned_pad/256] member of each edge workgroup
#pragma once
#include <hip/hip_runtime.h>

namespace dgr {

// info word bits
enum : int {
    EI_KNB_MASK = 0x3,      // local edge index of this edge in the neighbour
    EI_IS_IN = 1 << 2,      // this element is el_in (EdgtoEls[2e])
    EI_DIR_SAME = 1 << 3,   // el_in walks the edge in the same direction as this element
    EI_SGN_SHIFT = 4,       // sign(jac of el_in) + 1 in bits 4..5
    EI_PRESCRIBED = 1 << 6, // BdryPrescribed != 0
    EI_BOUNDARY = 1 << 7,   // el_in == el_ex
    EI_CODE_SHIFT = 8,      // BdryPrescribed code (clamped to 0..255) in bits 8..15
};

// node -> element CSR entry: element id and the node's local vertex index
enum : int {
    N2E_EL_MASK = 0x3fffffff,
    N2E_K_SHIFT = 30,
};

// the node kernels stage a 64-node group's element means in LDS when the
// group touches at most NL_CAP elements (a structured grid: 258)
constexpr int NL_CAP = 264;

// edge info word: bits 4..15 as in the element info word, plus
enum : int {
    ED_KIN_MASK = 0x3,      // local edge index in el_in
    ED_KEX_SHIFT = 2,       // local edge index in el_ex in bits 2..3
};

// Indices of the per-element geometry quantities that geometry_of() hands
// out (2DInnerProducts.c:22-308, compute2DL.c:98-144,380-391).
enum GeoIdx : int {
    GX_XE = 0,    // Gauss point x on edge e, point n: GX_XE + 2e + n
    GX_YE = 6,    // Gauss point y
    GX_PS = 12,   // CR basis b at Gauss point (e, n): GX_PS + 6b + 2e + n
    GX_NX = 30,   // midpoint-triangle normals nx1..nx3 (2DInnerProducts.c:111-116)
    GX_NY = 33,
    GX_M2 = 36,   // M12, M22, M32 of the RT0 fit (:126-128)
    GX_DEN = 39,  // Cramer determinant (:134)
    GX_EJ = 40,   // edge jacobians |e_k|/2 (:42-45)
    GX_GP = 43,   // grad Psi_b, component j: GX_GP + 2b + j (:303-308)
    GX_GS = 49,   // gradPsi_b . (x1+x2+x3, y1+y2+y3)
    GX_AJ = 52,   // |jac| of the midpoint triangle
    GX_AW = 53,   // |jac| * weight
    GX_MI = 54,   // Minv = 6/|jac_vertex| (compute2DL.c:382)
    GX_MZ = 55,   // midpoint depths (:268-270)
    GX_GNN = 58,  // g * n_mid * n_mid (:318)
    GX_BGX = 61,  // g |jac| dz/dx weight (:535)
    GX_BGY = 62,  // g |jac| dz/dy weight
    GX_ENX = 63,  // canonical (el_in) unit normal of local edge k (compute2DL.c:131-132)
    GX_ENY = 66,
    GX_EZB = 69,  // z_edge (compute2DL.c:144)
    GX_EDN = 72,  // 1/(nx*ty - ny*tx) (compute2DL.c:159,187)
    GX_COUNT = 75
};

// per-member constants
struct MemberDev {
    int el_off, ne;           // elements [el_off, el_off + ne)
    int ed_off, ned;          // edges [ed_off, ed_off + ned)
    int eblk0, eblks;         // element workgroups of the member
    int dblk0, dblks;         // edge workgroups of the member
    int tide_code;
    int hyd_n[3], hyd_off[3];  // hydrograph of boundary code c+1: hyd_n points at
                               // MeshDev::hydro + hyd_off (t, then zeta_b, then qn_b)
    int sentinels;             // prescribed edges: bzeta < 0 / bQn > 1000 take the interior value
    double rain, cfl, min_edge, final_time;
    double tide_datum, tide_amp, tide_period, tide_phase;
    double infil_K, infil_P;  // Green-Ampt conductivity and psi * delta_theta (K = 0: off)
};

// the edge reverse's list of non-wet edges per workgroup, the wet edges'
// tide adjoints of the listed workgroups and the queue of those workgroups
// (wet/dry on meshes of ADJ_EDGE_SPLIT padded edges or more whose elements
// are nearly all wet when the reverse starts): blist[ned_pad], tb[ned_pad],
// queue[ned_pad / 256], two counters used alternately (zero before the first
// call)
constexpr int ADJ_EDGE_SPLIT = 1 << 21;
struct AdjEdgeQueue {
    int *blist;
    double *tb;
    int2 *queue;
    int *qcount;
    mutable int parity;
    bool split;   // two launches (else the one-launch kernel)
};

struct MeshDev {
    int pitch, nn, ned_pad, nmem;
    const double4 *nodes;
    const double *zn;   // nodes[i].z as a plane, for kernels that need only the bed
    const int *ev;
    const int *nb;
    const int *info;
    const int *eid;
    const int *n2e_off;
    const int *n2e;
    const int *nl_off, *nl, *loc_off;
    const unsigned *locw;
    const int4 *edges;
    const int *einfo;
    const double4 *egeo;
    const MemberDev *mem;
    const int *eblk_mem;
    const int *dblk_mem;
    const double *hydro;      // hydrograph tables of all members
    int wg;                   // workgroup size of k_rhs, k_node_minmax and k_limit (64, 128 or 256)
};

// one per member
struct Ctrl {
    double t;                 // time at the start of the next step
    double dt;                // dt of the next step
    double tide;              // tide elevation for the current step
    double infil;             // Green-Ampt mean infiltration rate over the current step
    unsigned long long maxlam[2];   // stage max |lambda| as raw bits (values >= 0)
    // earliest error in the reference's execution order (atomic min):
    // step << 34 | phase << 32 | id << 3 | code, phase = 2*stage + (0 flux, 1 PDop);
    // id is member-local; all ones when no error has been raised
    unsigned long long err_key;
    long long step;                 // steps started (k_step_begin / k_step_mark)
    // max |lambda| is the reference's left fold of max(a,b)=(a>b?a:b) over the
    // edge-ordered operands (compute2DL.c:232-238,289-292); a NaN operand resets
    // the fold.  nan_pos = last NaN operand position (-1: none); the fold's
    // value is then the max of the operands after it (after_bits, any_after).
    long long nan_pos[2];
    unsigned long long after_bits[2];
    int any_after[2];
    // boundary values of the codes with a hydrograph at the current step
    double hyd_z[3], hyd_q[3];
};

// Hydrograph value at `time`: piecewise linear through n points (t strictly
// increasing), held at the end values outside them.  The oracle's
// orc_hydro_eval is the same function operation for operation.
__host__ __device__ inline double hydro_eval(int n, const double *t, const double *v, double time)
{
    if (time <= t[0]) return v[0];
    if (time >= t[n - 1]) return v[n - 1];
    int lo = 0, hi = n - 1;          // t[lo] <= time < t[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (t[mid] <= time) lo = mid;
        else hi = mid;
    }
    return v[lo] + (v[hi] - v[lo]) * ((time - t[lo]) / (t[hi] - t[lo]));
}

__host__ __device__ inline double maxlam_of(const Ctrl &c, int slot)
{
    unsigned long long b;
    if (c.nan_pos[slot] < 0) b = c.maxlam[slot];
    else if (c.any_after[slot]) b = c.after_bits[slot];
    else b = 0x7ff8000000000000ULL;
    double d;
    __builtin_memcpy(&d, &b, 8);
    return d;
}

__host__ __device__ inline int err_code_of(unsigned long long key) { return key == ~0ULL ? 0 : (int)(key & 7); }
__host__ __device__ inline int err_where_of(unsigned long long key) { return key == ~0ULL ? 0 : (int)((key >> 3) & 0x1fffffff); }

struct PhysDev {
    double g, h0, velzero;
    int wetdry;
    // per-workgroup max |lambda| partials of the flux kernels, [2][maxparts]
    // raw bits; reduced per member into Ctrl::maxlam by k_step_end /
    // k_lam_reduce (one plain store per workgroup, no same-address atomics)
    unsigned long long *lamparts;
    // edge kernel only, [2][maxparts]: member-local position of the last NaN
    // operand in the workgroup's edge range (-1: none); lamparts then holds
    // the max of the operands after it (-1.0: none)
    int *nanparts;
    // edge kernel of a taped stage 2 only, [maxparts]: the workgroup's
    // largest non-NaN operand (raw bits, -1.0: none) and how many of its
    // operands equal it; k_step_end turns them into StepRec::ntie
    unsigned long long *tiemax;
    int *tiecnt;
    int maxparts;
};

// per-member bookkeeping of the reverse sweep
struct AdjCtrl {
    double S;        // sum of the adjoints of the step start times after this step
    double lamb;     // adjoint of the max |lambda| that set the dt of the step being reversed
    double loss;
    int ntie;        // edges whose |lambda| operand equals the recorded max
    int unsupported; // set if the forward took a branch the reverse does not cover
};

// the reverse's per-step scalar pass (adj_scalar_member) for one step,
// carried by the next stage-2 limiter reverse launch (active = 0: none)
struct AdjScalarArgs {
    const double *dtparts;          // [2][maxparts] dt adjoint partials of the step's RHS reverses
    const double *tideparts;        // [2][maxparts] tide adjoint partials of its edge reverses
    const struct StepRec *rec;      // the step's record
    const struct StepRec *rec_prev; // the step before's
    int maxparts, has_prev, rec_ties, active;
};

// per-step record of the forward pass, per member
struct StepRec {
    double t, dt, tide;         // at the step's start
    unsigned long long lam2;    // max |lambda| of the step's second stage (raw bits)
    // operands of the stage-2 max-|lambda| fold equal to lam2 (taped edge-kernel
    // steps; -1: the fold saw a NaN operand)
    int ntie;
    int pad;
};

// which workgroup space a flux launch wrote its partials in
enum PartKind : int { PARTS_ELEM = 0, PARTS_EDGE = 1 };

// buffers of one persistent stage launch (k_stage_persist)
struct PersistArgs {
    const double *win;     // stage input
    double *u;             // stage output before the limiter (RK update)
    const double *wbase;   // step-start state (stage 2)
    const double *bzeta, *bqn;
    const int *wd;         // wet/dry flags the stage uses
    double *fhat, *means, *nodemm;
    double *lout;          // limiter (+ PDop) output
    unsigned char *modify;
    Ctrl *ctrls;
    const int *wd_lin;     // the limiter's status update: flags in and out (null: none)
    int *wd_lout;
    int next;              // stage 2: the step end does the next step's bookkeeping
};

}  // namespace dgr
