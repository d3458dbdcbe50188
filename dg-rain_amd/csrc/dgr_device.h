// dgr_device.h -- device-side data layout shared by the kernels and the host
// code of libdgrain.
//
// HBM layout (one instance):
//   state        double [9][pitch]   planes zeta0..2, qx0..2, qy0..2 (plane f*3+k
//                                     holds the reference's field f at el*3+k)
//   nodes        double4[num_nodes]  {x, y, z, n}: one 32-byte gather per vertex
//   ev           int    [3][pitch]   EltoVert, SoA
//   nb           int    [3][pitch]   neighbour element across local edge k
//                                     (the element itself on a boundary edge)
//   info         int    [3][pitch]   packed edge facts, see EI_* below
//   eid          int    [3][pitch]   global edge id (boundary data lookup)
//   means        double [3][pitch]   element averages of the updated state
//   nodemm       double [6][num_nodes]  node min (zeta,qx,qy) then max
//   n2e_off/n2e  node -> element CSR in ascending element order
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

struct MeshDev {
    int ne, nn, pitch;
    const double4 *nodes;
    const int *ev;
    const int *nb;
    const int *info;
    const int *eid;
    const int *n2e_off;
    const int *n2e;
};

struct Ctrl {
    double t;                 // time at the start of the next step
    double dt;                // dt of the next step
    double tide;              // tide elevation for the current step
    double pad0;
    unsigned long long maxlam[2];   // stage max |lambda| as raw bits (values >= 0)
    int err_code, err_where;
    int pad1[2];
};

struct PhysDev {
    double g, h0, velzero, rain;
    double cfl, min_edge, final_time;
    double tide_datum, tide_amp, tide_period, tide_phase;
    int wetdry, tide_code;
};

}  // namespace dgr
