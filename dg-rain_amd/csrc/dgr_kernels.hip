// dgr_kernels.hip -- gfx950 kernels of the 2D p=1 DG shallow-water RK step.
//
// One RK stage is four element/node-parallel phases:
//   k_wd_status   wetDryStatus2D            (2DCode/wetDry2D.c:16-60), wet/dry runs only
//   k_stage       compute2DL + RK update    (compute2DL.c:12-413, numericalFlux2D.c:6-108,
//                                            2DInnerProducts.c:8-570, time_evolution.c:72-88,208-217)
//   k_node_minmax node min/max of element means (SlopeLimiter2D.c:106-153)
//   k_limit       SlopeLimiter element pass (SlopeLimiter2D.c:157-343) + PDop2D (wetDry2D.c:63-260)
//
// The reference's edge loop (compute2DL.c:17-331) scatters each Roe flux to
// its two elements.  Here every element gathers its three fluxes instead: it
// evaluates each edge in the reference's canonical orientation (el_in's
// outward normal, el_in's state as the "in" state) and negates on the el_ex
// side, so every Fhat slot gets the reference's bits, with no atomics and no
// dependence on launch order.  Interior fluxes are therefore evaluated twice.
//
// Floating point: the kernels are compiled with -ffp-contract=off and keep the
// reference's evaluation order operation for operation; the only intended
// differences from the reference binary are pow(x,2) -> x*x and
// pow(H,4./3) -> dgr_pow43 (see dgr_math.h).
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dgrain.h"
#include "dgr_adjoint.h"
#include "dgr_device.h"
#include "dgr_math.h"

namespace dgr {

#define DGR_MAX(a, b) (((a) > (b)) ? (a) : (b))
#define DGR_MIN(a, b) (((a) < (b)) ? (a) : (b))

__device__ __forceinline__ int isign(double a) { return a < 0 ? -1 : (a == 0 ? 0 : 1); }

// Logical workgroup index: the dispatcher hands workgroup b to XCD b % 8;
// this bijection gives each XCD a contiguous range of workgroups (elements or
// edges), so neighbour gathers and partial-line stores of one region meet in
// one L2 instead of eight.
__device__ __forceinline__ int xcd_map(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7, j = b >> 3;
    return (x < r) ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}
__device__ __forceinline__ int xcd_block() { return xcd_map((int)blockIdx.x, (int)gridDim.x); }

// record an error; the smallest key (earliest step, stage, phase, then
// lowest member-local edge/element id) wins, which is the error the
// reference's serial loops would have stopped at (compute2DL.c:245-302,
// wetDry2D.c:142-146,233-237).  step_adj: steps by which Ctrl::step is
// already ahead of the step being computed (1 in a stage-2 limiter launch
// that follows a step end folded into the node bounds)
__device__ __forceinline__ void raise_error(Ctrl *c, int code, int where, int phase, int step_adj = 0)
{
    const unsigned long long key = ((unsigned long long)(c->step - step_adj) << 34) | ((unsigned long long)phase << 32) |
                                   ((unsigned long long)(where & 0x1fffffff) << 3) | (unsigned long long)code;
    atomicMin(&c->err_key, key);
}

// Prescribed boundary data of an edge with BdryPrescribed code `code`: the
// tide on the tide code, the current hydrograph values on a code that has
// one, else the edge's bzeta/bqn (boundary_conditions.c:10-83).
__device__ __forceinline__ void prescribed_values(const MemberDev &M, const Ctrl *ctrl, int code,
                                                  const double *bzeta, const double *bqn, int ed, double &zb,
                                                  double &qn)
{
    const bool tide = M.tide_code != 0 && code == M.tide_code;
    const int ci = (code >= 1 && code <= 3) ? code - 1 : 0;
    const bool hyd = code >= 1 && code <= 3 && M.hyd_n[ci] != 0;
    zb = tide ? ctrl->tide : hyd ? ctrl->hyd_z[ci] : bzeta[ed];
    qn = tide ? 1000000.0 : hyd ? ctrl->hyd_q[ci] : bqn[ed];
}

// ---------------------------------------------------------------------------
// Roe flux, numericalFlux2D.c:6-108.  Returns max |lambda| over the six
// one-sided eigenvalues (:101-105).
// ---------------------------------------------------------------------------
template <class DV>
__device__ __forceinline__ double roe_flux(double F[3], double zi, double ze, double qxi, double qxe,
                                           double qyi, double qye, double zb, double nx, double ny,
                                           double lg, double g, DV &D)
{
    const double Hi = zi + zb, He = ze + zb;
    const typename DV::Den DHi = D.den(Hi), DHe = D.den(He);
    const double ui = D.q(qxi, DHi), ue = D.q(qxe, DHe);
    const double vi = D.q(qyi, DHi), ve = D.q(qye, DHe);
    const double j0 = zi - ze, j1 = qxi - qxe, j2 = qyi - qye;
    const double f1xi = Hi * ui, f1xe = He * ue;
    const double f1yi = Hi * vi, f1ye = He * ve;
    const double f2xi = Hi * ui * ui + 0.5 * lg * (Hi * Hi - zb * zb);
    const double f2xe = He * ue * ue + 0.5 * lg * (He * He - zb * zb);
    const double f2yi = Hi * ui * vi, f2ye = He * ue * ve;
    const double f3xi = Hi * ui * vi, f3xe = He * ue * ve;
    const double f3yi = Hi * vi * vi + 0.5 * lg * (Hi * Hi - zb * zb);
    const double f3ye = He * ve * ve + 0.5 * lg * (He * He - zb * zb);
    const double fni0 = f1xi * nx + f1yi * ny, fne0 = f1xe * nx + f1ye * ny;
    const double fni1 = f2xi * nx + f2yi * ny, fne1 = f2xe * nx + f2ye * ny;
    const double fni2 = f3xi * nx + f3yi * ny, fne2 = f3xe * nx + f3ye * ny;

    const double Hh = 0.5 * (Hi + He);
    const double sHi = sqrt(Hi), sHe = sqrt(He);
    const typename DV::Den DS = D.den(sHi + sHe);
    const double uh = D.q(ui * sHi + ue * sHe, DS);
    const double vh = D.q(vi * sHi + ve * sHe, DS);
    const double c = sqrt(g * Hh);

    const double R10 = uh - c * nx, R11 = -ny, R12 = uh + c * nx;
    const double R20 = vh - c * ny, R21 = nx, R22 = vh + c * ny;
    const typename DV::Den den = D.den(-2 * c);
    const double Q00 = D.q(-c - nx * uh - ny * vh, den);
    const double Q01 = D.q(nx, den);
    const double Q02 = D.q(ny, den);
    const double Q10 = D.q(-2 * c * ny * uh + 2 * c * nx * vh, den);
    const double Q11 = D.q(2 * c * ny, den);
    const double Q12 = D.q(-2 * c * nx, den);
    const double Q20 = D.q(-c + nx * uh + ny * vh, den);
    // (-nx)/den is -(nx/den): rounding to nearest is symmetric
    const double Q21 = -Q01;
    const double Q22 = -Q02;

    const double sgi = sqrt(g * Hi), sge = sqrt(g * He);
    const double l1i = ui * nx + vi * ny - sgi;
    const double l2i = ui * nx + vi * ny;
    const double l3i = ui * nx + vi * ny + sgi;
    const double l1e = ue * nx + ve * ny - sge;
    const double l2e = ue * nx + ve * ny;
    const double l3e = ue * nx + ve * ny + sge;
    double lh1 = uh * nx + vh * ny - c;
    double lh2 = uh * nx + vh * ny;
    double lh3 = uh * nx + vh * ny + c;

    double e1 = fmax(0.0, (lh1 - l1i));
    e1 = fmax(e1, (l1e - lh1));
    lh1 = fmax(e1, fabs(lh1));
    double e2 = fmax(0.0, (lh2 - l2i));
    e2 = fmax(e2, (l2e - lh2));
    lh2 = fmax(e2, fabs(lh2));
    double e3 = fmax(0.0, (lh3 - l3i));
    e3 = fmax(e3, (l3e - lh3));
    lh3 = fmax(e3, fabs(lh3));

    // MatrixVectorMultiply x3 (mathfunctions.c:24-35), zeros included
    const double p10 = ((0.0 + Q00 * j0) + Q01 * j1) + Q02 * j2;
    const double p11 = ((0.0 + Q10 * j0) + Q11 * j1) + Q12 * j2;
    const double p12 = ((0.0 + Q20 * j0) + Q21 * j1) + Q22 * j2;
    const double p20 = ((0.0 + lh1 * p10) + 0.0 * p11) + 0.0 * p12;
    const double p21 = ((0.0 + 0.0 * p10) + lh2 * p11) + 0.0 * p12;
    const double p22 = ((0.0 + 0.0 * p10) + 0.0 * p11) + lh3 * p12;
    const double p30 = ((0.0 + 1.0 * p20) + 0.0 * p21) + 1.0 * p22;
    const double p31 = ((0.0 + R10 * p20) + R11 * p21) + R12 * p22;
    const double p32 = ((0.0 + R20 * p20) + R21 * p21) + R22 * p22;
    F[0] = 0.5 * (fni0 + fne0) + 0.5 * p30;
    F[1] = 0.5 * (fni1 + fne1) + 0.5 * p31;
    F[2] = 0.5 * (fni2 + fne2) + 0.5 * p32;

    double mx = DGR_MAX(fabs(l1i), fabs(l1e));
    mx = DGR_MAX(mx, fabs(l2i));
    mx = DGR_MAX(mx, fabs(l2e));
    mx = DGR_MAX(mx, fabs(l3i));
    mx = DGR_MAX(mx, fabs(l3e));
    return mx;
}

// Cramer solve of the RT0 fit (2DInnerProducts.c:134-146)
template <class DV>
__device__ __forceinline__ void rt0(double M11, double M12, double M13, double M21, double M22, double M23,
                                    double M31, double M32, double M33, const typename DV::Den &den, double R1,
                                    double R2, double R3, double &a, double &b, double &c, DV &D)
{
    a = D.q(R3 * M13 * M22 - R3 * M12 * M23 - R2 * M13 * M32 + R1 * M23 * M32 + R2 * M12 * M33 - R1 * M22 * M33, den);
    b = D.q(-(R3 * M13 * M21 - R3 * M11 * M23 - R2 * M13 * M31 + R1 * M23 * M31 + R2 * M11 * M33 - R1 * M21 * M33), den);
    c = D.q(R3 * M12 * M21 - R3 * M11 * M22 - R2 * M12 * M31 + R1 * M22 * M31 + R2 * M11 * M32 - R1 * M21 * M32, den);
}

// ---------------------------------------------------------------------------
// geometry_of: every geometry-only quantity of Compute2DInnerProducts and of
// the edge loop of compute2DL for element e, with the reference's
// expressions, handed to put(GeoIdx, value).
// ---------------------------------------------------------------------------
template <typename Put, class DV>
__device__ __forceinline__ void geometry_of_vertices(const MeshDev &m, int e, const double4 P[3], double g, Put &&put,
                                                     DV &D)
{
    const int Pt = m.pitch;
    const double xv1 = P[0].x, xv2 = P[1].x, xv3 = P[2].x;                  // 2DInnerProducts.c:22-27
    const double yv1 = P[0].y, yv2 = P[1].y, yv3 = P[2].y;
    const double x1 = (xv1 + xv2) / 2, x2 = (xv2 + xv3) / 2, x3 = (xv3 + xv1) / 2;   // :30-35
    const double y1 = (yv1 + yv2) / 2, y2 = (yv2 + yv3) / 2, y3 = (yv3 + yv1) / 2;
    const double el0 = sqrt((xv2 - xv1) * (xv2 - xv1) + (yv2 - yv1) * (yv2 - yv1));   // :37-39
    const double el1 = sqrt((xv3 - xv2) * (xv3 - xv2) + (yv3 - yv2) * (yv3 - yv2));
    const double el2 = sqrt((xv3 - xv1) * (xv3 - xv1) + (yv3 - yv1) * (yv3 - yv1));
    put(GX_EJ + 0, el0 / 2);
    put(GX_EJ + 1, el1 / 2);
    put(GX_EJ + 2, el2 / 2);

    const double sqrt3 = sqrt(3.0);                                          // :51-65
    const typename DV::Den D3 = D.den_const(sqrt3);
    double xe[3][2], ye[3][2];
    xe[0][0] = D.q(sqrt3 * x1 - x2 + x3, D3);  ye[0][0] = D.q(sqrt3 * y1 - y2 + y3, D3);
    xe[0][1] = D.q(sqrt3 * x1 + x2 - x3, D3);  ye[0][1] = D.q(sqrt3 * y1 + y2 - y3, D3);
    xe[1][0] = D.q(x1 + sqrt3 * x2 - x3, D3);  ye[1][0] = D.q(y1 + sqrt3 * y2 - y3, D3);
    xe[1][1] = D.q(-x1 + sqrt3 * x2 + x3, D3); ye[1][1] = D.q(-y1 + sqrt3 * y2 + y3, D3);
    xe[2][0] = D.q(-x1 + x2 + sqrt3 * x3, D3); ye[2][0] = D.q(-y1 + y2 + sqrt3 * y3, D3);
    xe[2][1] = D.q(x1 - x2 + sqrt3 * x3, D3);  ye[2][1] = D.q(y1 - y2 + sqrt3 * y3, D3);
    const double dn0 = x1 * y2 - x1 * y3 + x2 * y3 - x2 * y1 + x3 * y1 - x3 * y2;   // :73
    const typename DV::Den Ddn = D.den(dn0);
#pragma unroll
    for (int ed = 0; ed < 3; ++ed)
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            const double xq = xe[ed][n], yq = ye[ed][n];
            put(GX_XE + 2 * ed + n, xq);
            put(GX_YE + 2 * ed + n, yq);
            put(GX_PS + 0 + 2 * ed + n, D.q(-(yq * x2 - yq * x3 - xq * y2 + x3 * y2 + xq * y3 - x2 * y3), Ddn));   // :74-87
            put(GX_PS + 6 + 2 * ed + n, D.q(-(-x1 * yq + x3 * yq - x3 * y1 + xq * y1 - xq * y3 + x1 * y3), Ddn));
            put(GX_PS + 12 + 2 * ed + n, D.q(-(yq * x1 - yq * x2 - xq * y1 + x2 * y1 + xq * y2 - x1 * y2), Ddn));
        }
    const double jac = 4 * (x2 - x3) * (y2 - y1) - 4 * (x2 - x1) * (y2 - y3);  // :110
    const int sj = isign(jac);
    const typename DV::Den De0 = D.den(el0), De1 = D.den(el1), De2 = D.den(el2);   // :111-116
    const double nx1 = D.q(sj * 2 * (y2 - y3), De0), ny1 = D.q(sj * 2 * (x3 - x2), De0);
    const double nx2 = D.q(sj * 2 * (y3 - y1), De1), ny2 = D.q(sj * 2 * (x1 - x3), De1);
    const double nx3 = D.q(sj * 2 * (y1 - y2), De2), ny3 = D.q(sj * 2 * (x2 - x1), De2);
    const double M11 = nx1, M12 = x1 * nx1 + y1 * ny1, M13 = ny1;            // :126-128
    const double M21 = nx2, M22 = x2 * nx2 + y2 * ny2, M23 = ny2;
    const double M31 = nx3, M32 = x3 * nx3 + y3 * ny3, M33 = ny3;
    put(GX_NX + 0, nx1); put(GX_NX + 1, nx2); put(GX_NX + 2, nx3);
    put(GX_NY + 0, ny1); put(GX_NY + 1, ny2); put(GX_NY + 2, ny3);
    put(GX_M2 + 0, M12); put(GX_M2 + 1, M22); put(GX_M2 + 2, M32);
    put(GX_DEN, M13 * M22 * M31 - M12 * M23 * M31 - M13 * M21 * M32 + M11 * M23 * M32 + M12 * M21 * M33 - M11 * M22 * M33);

    const double weight = 1.0 / 6;                                            // :247
    const double mz[3] = {0.5 * (P[0].z + P[1].z), 0.5 * (P[1].z + P[2].z), 0.5 * (P[2].z + P[0].z)};
    const double mn[3] = {0.5 * (P[0].w + P[1].w), 0.5 * (P[1].w + P[2].w), 0.5 * (P[2].w + P[0].w)};
    const double fj = D.q(4.0, D.den(jac));                                  // 4 / jac
    const double gP[3][2] = {{fj * (y2 - y3), fj * (x3 - x2)},               // :303-308
                             {fj * (y3 - y1), fj * (x1 - x3)},
                             {fj * (y1 - y2), fj * (x2 - x1)}};
    const double sx = x1 + x2 + x3, sy = y1 + y2 + y3;
    const double aj = fabs(jac);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        put(GX_GP + 2 * b, gP[b][0]);
        put(GX_GP + 2 * b + 1, gP[b][1]);
        put(GX_GS + b, gP[b][0] * sx + gP[b][1] * sy);
        put(GX_MZ + b, mz[b]);
        put(GX_GNN + b, g * mn[b] * mn[b]);
    }
    put(GX_AJ, aj);
    put(GX_AW, aj * weight);
    const double gz0 = mz[0] * gP[0][0] + mz[1] * gP[1][0] + mz[2] * gP[2][0];   // :343-347
    const double gz1 = mz[0] * gP[0][1] + mz[1] * gP[1][1] + mz[2] * gP[2][1];
    put(GX_BGX, g * aj * gz0 * weight);                                       // :535
    put(GX_BGY, g * aj * gz1 * weight);
    const double jv = P[0].x * (P[1].y - P[2].y) + P[1].x * (P[2].y - P[0].y) + P[2].x * (P[0].y - P[1].y);
    put(GX_MI, D.q(6.0, D.den(fabs(jv))));                                    // compute2DL.c:380-382

    // edge loop geometry in the canonical orientation (compute2DL.c:98-144)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int inf = m.info[k * Pt + e];
        const double4 A = P[k], B = P[(k + 1) % 3];
        const double len = sqrt((B.x - A.x) * (B.x - A.x) + (B.y - A.y) * (B.y - A.y));
        const double sg = (double)(((inf >> EI_SGN_SHIFT) & 3) - 1);
        double x0, y0, xx1, yy1;
        if (inf & EI_DIR_SAME) { x0 = A.x; y0 = A.y; xx1 = B.x; yy1 = B.y; }
        else { x0 = B.x; y0 = B.y; xx1 = A.x; yy1 = A.y; }
        const typename DV::Den Dl = D.den(len);
        const double nx = D.q(sg * (yy1 - y0), Dl);
        const double ny = D.q(sg * (x0 - xx1), Dl);
        const double tx = -ny, ty = nx;
        put(GX_ENX + k, nx);
        put(GX_ENY + k, ny);
        put(GX_EZB + k, (A.z + B.z) / 2);
        put(GX_EDN + k, D.q(1.0, D.den(nx * ty - ny * tx)));
    }
}

template <typename Put, class DV>
__device__ __forceinline__ void geometry_of(const MeshDev &m, int e, double g, Put &&put, DV &D)
{
    double4 P[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) P[k] = m.nodes[m.ev[k * m.pitch + e]];
    geometry_of_vertices(m, e, P, g, put, D);
}

template <typename Put>
__device__ __forceinline__ void geometry_of(const MeshDev &m, int e, double g, Put &&put)
{
    DivExact D;
    geometry_of(m, e, g, put, D);
}

// geometry computed into registers; dead-code elimination keeps only what
// the caller reads
struct GeoRegs {
    double v[GX_COUNT];
    __device__ __forceinline__ double operator()(int i) const { return v[i]; }
};

// ---------------------------------------------------------------------------
// Compute2DInnerProducts + the element loop of compute2DL: the 9 RHS values
// of one element from its state W[f*3+k], its fluxes Fh[f*3+k] and its
// geometry planes.  2DInnerProducts.c:14-566, compute2DL.c:344-409.
// ---------------------------------------------------------------------------
template <typename Geo, class DV>
__device__ __forceinline__ void element_rhs(const Geo &G, const double W[9], const double Fh[9], int wd_self,
                                            double g, double rain, double infil, double R[9], DV &D)
{
    const double nx1 = G(GX_NX), nx2 = G(GX_NX + 1), nx3 = G(GX_NX + 2);
    const double ny1 = G(GX_NY), ny2 = G(GX_NY + 1), ny3 = G(GX_NY + 2);
    const double M11 = nx1, M12 = G(GX_M2), M13 = ny1;
    const double M21 = nx2, M22 = G(GX_M2 + 1), M23 = ny2;
    const double M31 = nx3, M32 = G(GX_M2 + 2), M33 = ny3;
    const typename DV::Den den = D.den(G(GX_DEN));
    const double nxs[3] = {nx1, nx2, nx3}, nys[3] = {ny1, ny2, ny3};

    // surface integrals SI[f*3+b] (:134-239)
    double a[3], b[3], c[3];
#pragma unroll
    for (int f = 0; f < 3; ++f)
        rt0(M11, M12, M13, M21, M22, M23, M31, M32, M33, den, Fh[f * 3 + 0], Fh[f * 3 + 1], Fh[f * 3 + 2],
            a[f], b[f], c[f], D);
    double SI[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) SI[f] = 0.0;
#pragma unroll
    for (int ed = 0; ed < 3; ++ed) {
        const double ej = G(GX_EJ + ed);
        double Fn[3][2];
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            const double xq = G(GX_XE + 2 * ed + n), yq = G(GX_YE + 2 * ed + n);
#pragma unroll
            for (int f = 0; f < 3; ++f) Fn[f][n] = (a[f] + b[f] * xq) * nxs[ed] + (c[f] + b[f] * yq) * nys[ed];
        }
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) {
            const double p0 = G(GX_PS + 6 * bb + 2 * ed), p1 = G(GX_PS + 6 * bb + 2 * ed + 1);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const double term = ej * (Fn[f][0] * p0 + Fn[f][1] * p1);
                // ((term_e0 + term_e1) + term_e2), the reference's left-to-right sum
                SI[f * 3 + bb] = (ed == 0) ? term : SI[f * 3 + bb] + term;
            }
        }
    }

    // volume terms (:247-566)
    const double weight = 1.0 / 6;
    const double lg = (wd_self == 0) ? 0.0 : g;
    double H[3], tau[3], Fx[3][3], Fy[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double mz = G(GX_MZ + k);
        const double zr = W[k], qx = W[3 + k], qy = W[6 + k];
        H[k] = zr + mz;
        const typename DV::Den DH = D.den(H[k]);
        const double u = D.q(qx, DH);
        const double v = D.q(qy, DH);
        tau[k] = D.q(G(GX_GNN + k) * sqrt(u * u + v * v), D.den(dgr_pow43(H[k])));
        const double uv = D.q(qx * qy, DH);
        Fx[0][k] = qx;
        Fx[1][k] = D.q(qx * qx, DH) + 0.5 * lg * ((H[k] * H[k]) - (mz * mz));
        Fx[2][k] = uv;
        Fy[0][k] = qy;
        Fy[1][k] = uv;
        Fy[2][k] = D.q(qy * qy, DH) + 0.5 * lg * ((H[k] * H[k]) - (mz * mz));
    }
    const double aj = G(GX_AJ);
    double VF[3][3];   // [field][basis]
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        double va, vb, vc;
        rt0(M11, M12, M13, M21, M22, M23, M31, M32, M33, den,
            Fx[f][0] * nx1 + Fy[f][0] * ny1, Fx[f][1] * nx2 + Fy[f][1] * ny2, Fx[f][2] * nx3 + Fy[f][2] * ny3,
            va, vb, vc, D);
#pragma unroll
        for (int bb = 0; bb < 3; ++bb)
            VF[f][bb] = aj * (0.5 * va * G(GX_GP + 2 * bb) + 0.5 * vc * G(GX_GP + 2 * bb + 1) + vb * weight * G(GX_GS + bb));
    }
    // rhs combine (compute2DL.c:344-354) and Minv (:380-404)
    const double Mi = G(GX_MI), aw = G(GX_AW), bgx = G(GX_BGX), bgy = G(GX_BGY);
    double ze[3], qe[3], re[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ze[k] = -SI[k] + VF[0][k];
        // with lg = 0 the bed term is (0*|jac|*dz*weight)*zeta = +-0 (a NaN zeta stays NaN)
        const double bx = (wd_self == 0) ? 0.0 * W[k] : bgx * W[k];
        const double fx = aw * tau[k] * W[3 + k];
        qe[k] = -SI[3 + k] + VF[1][k] + bx - fx;
        const double by = (wd_self == 0) ? 0.0 * W[k] : bgy * W[k];
        const double fy = aw * tau[k] * W[6 + k];
        re[k] = -SI[6 + k] + VF[2][k] + by - fy;
    }
    R[0] = (Mi * ze[0] + 0.0 * ze[1]) + 0.0 * ze[2];
    R[1] = (0.0 * ze[0] + Mi * ze[1]) + 0.0 * ze[2];
    R[2] = (0.0 * ze[0] + 0.0 * ze[1]) + Mi * ze[2];
    R[3] = (Mi * qe[0] + 0.0 * qe[1]) + 0.0 * qe[2];
    R[4] = (0.0 * qe[0] + Mi * qe[1]) + 0.0 * qe[2];
    R[5] = (0.0 * qe[0] + 0.0 * qe[1]) + Mi * qe[2];
    R[6] = (Mi * re[0] + 0.0 * re[1]) + 0.0 * re[2];
    R[7] = (0.0 * re[0] + Mi * re[1]) + 0.0 * re[2];
    R[8] = (0.0 * re[0] + 0.0 * re[1]) + Mi * re[2];
    if (rain != 0.0) {
        R[0] = R[0] + rain;
        R[1] = R[1] + rain;
        R[2] = R[2] + rain;
    }
    // infiltration leaves elements not flagged dry
    if (infil != 0.0 && wd_self != 0) {
        R[0] = R[0] - infil;
        R[1] = R[1] - infil;
        R[2] = R[2] - infil;
    }
}

// ---------------------------------------------------------------------------
// The same RHS as geometry_of + element_rhs, each geometry quantity computed
// with the same expression but next to its first use, in three phases:
// surface integrals (Gauss points and basis values one edge at a time),
// volume terms, combine.  Only the vertex data, the RT0 matrix and the
// running sums stay live across phases, instead of all 63 geometry values:
// the RHS kernel then fits fewer registers (more waves per SIMD).
// 2DInnerProducts.c:14-566, compute2DL.c:344-409.
// ---------------------------------------------------------------------------
template <class DV>
__device__ __forceinline__ void element_rhs_phased(const double4 P[3], const double W[9], const double Fh[9],
                                                   int wd_self, double g, double rain, double infil, double R[9],
                                                   DV &D)
{
    const double xv1 = P[0].x, xv2 = P[1].x, xv3 = P[2].x;                  // 2DInnerProducts.c:22-27
    const double yv1 = P[0].y, yv2 = P[1].y, yv3 = P[2].y;
    const double x1 = (xv1 + xv2) / 2, x2 = (xv2 + xv3) / 2, x3 = (xv3 + xv1) / 2;   // :30-35
    const double y1 = (yv1 + yv2) / 2, y2 = (yv2 + yv3) / 2, y3 = (yv3 + yv1) / 2;
    const double el0 = sqrt((xv2 - xv1) * (xv2 - xv1) + (yv2 - yv1) * (yv2 - yv1));   // :37-39
    const double el1 = sqrt((xv3 - xv2) * (xv3 - xv2) + (yv3 - yv2) * (yv3 - yv2));
    const double el2 = sqrt((xv3 - xv1) * (xv3 - xv1) + (yv3 - yv1) * (yv3 - yv1));
    const double jac = 4 * (x2 - x3) * (y2 - y1) - 4 * (x2 - x1) * (y2 - y3);  // :110
    const int sj = isign(jac);
    double nx1, ny1, nx2, ny2, nx3, ny3;
    {
        const typename DV::Den De0 = D.den(el0), De1 = D.den(el1), De2 = D.den(el2);   // :111-116
        nx1 = D.q(sj * 2 * (y2 - y3), De0); ny1 = D.q(sj * 2 * (x3 - x2), De0);
        nx2 = D.q(sj * 2 * (y3 - y1), De1); ny2 = D.q(sj * 2 * (x1 - x3), De1);
        nx3 = D.q(sj * 2 * (y1 - y2), De2); ny3 = D.q(sj * 2 * (x2 - x1), De2);
    }
    const double M11 = nx1, M12 = x1 * nx1 + y1 * ny1, M13 = ny1;            // :126-128
    const double M21 = nx2, M22 = x2 * nx2 + y2 * ny2, M23 = ny2;
    const double M31 = nx3, M32 = x3 * nx3 + y3 * ny3, M33 = ny3;
    const typename DV::Den den =
        D.den(M13 * M22 * M31 - M12 * M23 * M31 - M13 * M21 * M32 + M11 * M23 * M32 + M12 * M21 * M33 - M11 * M22 * M33);
    const double nxs[3] = {nx1, nx2, nx3}, nys[3] = {ny1, ny2, ny3};

    // surface integrals SI[f*3+b] (:134-239), Gauss points (:51-65) and
    // basis values (:73-87) of one edge at a time
    double SI[9];
    {
        double a[3], b[3], c[3];
#pragma unroll
        for (int f = 0; f < 3; ++f)
            rt0(M11, M12, M13, M21, M22, M23, M31, M32, M33, den, Fh[f * 3 + 0], Fh[f * 3 + 1], Fh[f * 3 + 2],
                a[f], b[f], c[f], D);
        const double sqrt3 = sqrt(3.0);
        const typename DV::Den D3 = D.den_const(sqrt3);
        const double dn0 = x1 * y2 - x1 * y3 + x2 * y3 - x2 * y1 + x3 * y1 - x3 * y2;   // :73
        const typename DV::Den Ddn = D.den(dn0);
        const double els[3] = {el0, el1, el2};
#pragma unroll
        for (int ed = 0; ed < 3; ++ed) {
            const double ej = els[ed] / 2;
            double xe[2], ye[2];
            if (ed == 0) {
                xe[0] = D.q(sqrt3 * x1 - x2 + x3, D3);  ye[0] = D.q(sqrt3 * y1 - y2 + y3, D3);
                xe[1] = D.q(sqrt3 * x1 + x2 - x3, D3);  ye[1] = D.q(sqrt3 * y1 + y2 - y3, D3);
            } else if (ed == 1) {
                xe[0] = D.q(x1 + sqrt3 * x2 - x3, D3);  ye[0] = D.q(y1 + sqrt3 * y2 - y3, D3);
                xe[1] = D.q(-x1 + sqrt3 * x2 + x3, D3); ye[1] = D.q(-y1 + sqrt3 * y2 + y3, D3);
            } else {
                xe[0] = D.q(-x1 + x2 + sqrt3 * x3, D3); ye[0] = D.q(-y1 + y2 + sqrt3 * y3, D3);
                xe[1] = D.q(x1 - x2 + sqrt3 * x3, D3);  ye[1] = D.q(y1 - y2 + sqrt3 * y3, D3);
            }
            double Fn[3][2], ps[3][2];
#pragma unroll
            for (int n = 0; n < 2; ++n) {
                const double xq = xe[n], yq = ye[n];
#pragma unroll
                for (int f = 0; f < 3; ++f) Fn[f][n] = (a[f] + b[f] * xq) * nxs[ed] + (c[f] + b[f] * yq) * nys[ed];
                ps[0][n] = D.q(-(yq * x2 - yq * x3 - xq * y2 + x3 * y2 + xq * y3 - x2 * y3), Ddn);
                ps[1][n] = D.q(-(-x1 * yq + x3 * yq - x3 * y1 + xq * y1 - xq * y3 + x1 * y3), Ddn);
                ps[2][n] = D.q(-(yq * x1 - yq * x2 - xq * y1 + x2 * y1 + xq * y2 - x1 * y2), Ddn);
            }
#pragma unroll
            for (int bb = 0; bb < 3; ++bb) {
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const double term = ej * (Fn[f][0] * ps[bb][0] + Fn[f][1] * ps[bb][1]);
                    // ((term_e0 + term_e1) + term_e2), the reference's left-to-right sum
                    SI[f * 3 + bb] = (ed == 0) ? term : SI[f * 3 + bb] + term;
                }
            }
        }
    }

    // volume terms (:247-566)
    const double weight = 1.0 / 6;
    const double lg = (wd_self == 0) ? 0.0 : g;
    const double mz[3] = {0.5 * (P[0].z + P[1].z), 0.5 * (P[1].z + P[2].z), 0.5 * (P[2].z + P[0].z)};
    const double aj = fabs(jac);
    double tau[3], VF[3][3];   // VF[field][basis]
    double bgx, bgy;
    {
        double Fx[3][3], Fy[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double mnk = 0.5 * (P[k].w + P[(k + 1) % 3].w);
            const double gnn = g * mnk * mnk;
            const double zr = W[k], qx = W[3 + k], qy = W[6 + k];
            const double H = zr + mz[k];
            const typename DV::Den DH = D.den(H);
            const double u = D.q(qx, DH);
            const double v = D.q(qy, DH);
            tau[k] = D.q(gnn * sqrt(u * u + v * v), D.den(dgr_pow43(H)));
            const double uv = D.q(qx * qy, DH);
            Fx[0][k] = qx;
            Fx[1][k] = D.q(qx * qx, DH) + 0.5 * lg * ((H * H) - (mz[k] * mz[k]));
            Fx[2][k] = uv;
            Fy[0][k] = qy;
            Fy[1][k] = uv;
            Fy[2][k] = D.q(qy * qy, DH) + 0.5 * lg * ((H * H) - (mz[k] * mz[k]));
        }
        const double fj = D.q(4.0, D.den(jac));                              // 4 / jac
        const double gP[3][2] = {{fj * (y2 - y3), fj * (x3 - x2)},           // :303-308
                                 {fj * (y3 - y1), fj * (x1 - x3)},
                                 {fj * (y1 - y2), fj * (x2 - x1)}};
        const double sx = x1 + x2 + x3, sy = y1 + y2 + y3;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            double va, vb, vc;
            rt0(M11, M12, M13, M21, M22, M23, M31, M32, M33, den,
                Fx[f][0] * nx1 + Fy[f][0] * ny1, Fx[f][1] * nx2 + Fy[f][1] * ny2, Fx[f][2] * nx3 + Fy[f][2] * ny3,
                va, vb, vc, D);
#pragma unroll
            for (int bb = 0; bb < 3; ++bb) {
                const double gs = gP[bb][0] * sx + gP[bb][1] * sy;
                VF[f][bb] = aj * (0.5 * va * gP[bb][0] + 0.5 * vc * gP[bb][1] + vb * weight * gs);
            }
        }
        const double gz0 = mz[0] * gP[0][0] + mz[1] * gP[1][0] + mz[2] * gP[2][0];   // :343-347
        const double gz1 = mz[0] * gP[0][1] + mz[1] * gP[1][1] + mz[2] * gP[2][1];
        bgx = g * aj * gz0 * weight;                                          // :535
        bgy = g * aj * gz1 * weight;
    }

    // rhs combine (compute2DL.c:344-354) and Minv (:380-404)
    const double jv = P[0].x * (P[1].y - P[2].y) + P[1].x * (P[2].y - P[0].y) + P[2].x * (P[0].y - P[1].y);
    const double Mi = D.q(6.0, D.den(fabs(jv)));                              // compute2DL.c:380-382
    const double aw = aj * weight;
    double ze[3], qe[3], re[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ze[k] = -SI[k] + VF[0][k];
        // with lg = 0 the bed term is (0*|jac|*dz*weight)*zeta = +-0 (a NaN zeta stays NaN)
        const double bx = (wd_self == 0) ? 0.0 * W[k] : bgx * W[k];
        const double fx = aw * tau[k] * W[3 + k];
        qe[k] = -SI[3 + k] + VF[1][k] + bx - fx;
        const double by = (wd_self == 0) ? 0.0 * W[k] : bgy * W[k];
        const double fy = aw * tau[k] * W[6 + k];
        re[k] = -SI[6 + k] + VF[2][k] + by - fy;
    }
    R[0] = (Mi * ze[0] + 0.0 * ze[1]) + 0.0 * ze[2];
    R[1] = (0.0 * ze[0] + Mi * ze[1]) + 0.0 * ze[2];
    R[2] = (0.0 * ze[0] + 0.0 * ze[1]) + Mi * ze[2];
    R[3] = (Mi * qe[0] + 0.0 * qe[1]) + 0.0 * qe[2];
    R[4] = (0.0 * qe[0] + Mi * qe[1]) + 0.0 * qe[2];
    R[5] = (0.0 * qe[0] + 0.0 * qe[1]) + Mi * qe[2];
    R[6] = (Mi * re[0] + 0.0 * re[1]) + 0.0 * re[2];
    R[7] = (0.0 * re[0] + Mi * re[1]) + 0.0 * re[2];
    R[8] = (0.0 * re[0] + 0.0 * re[1]) + Mi * re[2];
    if (rain != 0.0) {
        R[0] = R[0] + rain;
        R[1] = R[1] + rain;
        R[2] = R[2] + rain;
    }
    // infiltration leaves elements not flagged dry
    if (infil != 0.0 && wd_self != 0) {
        R[0] = R[0] - infil;
        R[1] = R[1] - infil;
        R[2] = R[2] - infil;
    }
}

// wall reflection of (qx, qy) about the edge normal (compute2DL.c:179-189);
// dn = 1/(nx*ty - ny*tx) with (tx, ty) = (-ny, nx)
__device__ __forceinline__ void reflect(double qx, double qy, double nx, double ny, double dn, double &rx, double &ry)
{
    const double tx = -ny, ty = nx;
    const double qn = qx * nx + qy * ny;
    const double qt = qx * tx + qy * ty;
    const double qne = -qn, qte = qt;
    rx = (ty * qne - ny * qte) * dn;
    ry = (-tx * qne + nx * qte) * dn;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T w = __shfl_xor(v, o, 64);
        v = (w > v) ? w : v;
    }
    return v;
}

// One max-|lambda| operand of the reference's fold, at position 2*edge (the
// flux of a not-both-dry edge, or el_in's reflection flux of a both-dry edge)
// or 2*edge+1 (el_ex's reflection flux of a both-dry edge).
// Two fixed slots per local edge (2k, 2k+1) so every index is a compile-time
// constant after unrolling; pos < 0 marks an unused slot.
struct LamOps {
    double v[6];
    long long pos[6];
};

// The three Roe fluxes of element e in the reference's canonical orientation
// (compute2DL.c:17-331), negated on the el_ex side, and the max-|lambda|
// operands this element owns.
template <bool WETDRY, typename Geo>
__device__ __forceinline__ void element_fluxes(const MeshDev &m, const PhysDev &ph, const MemberDev &M,
                                               const double *__restrict__ win,
                                               const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                               const int *__restrict__ wd, const Ctrl *ctrl, int e,
                                               const Geo &G, const double W[9], int wd_self, double Fh[9],
                                               LamOps &ops, int &err_code, int &err_edge)
{
    const int P = m.pitch;
    const double g = ph.g;
    DivExact D;
#pragma unroll
    for (int i = 0; i < 6; ++i) { ops.v[i] = 0.0; ops.pos[i] = -1; }
    err_code = 0;
    err_edge = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int inf = m.info[k * P + e];
        const int nbk = m.nb[k * P + e];
        const int ed = m.eid[k * P + e];
        const int knb = inf & EI_KNB_MASK;
        const bool is_in = (inf & EI_IS_IN) != 0;
        const bool boundary = (inf & EI_BOUNDARY) != 0;
        const double nx = G(GX_ENX + k);
        const double ny = G(GX_ENY + k);
        const double zb = G(GX_EZB + k);
        const double dn = G(GX_EDN + k);
        const double tx = -ny, ty = nx;
        const double oz = W[k], oqx = W[3 + k], oqy = W[6 + k];
        double zin, qxin, qyin, zex, qxex, qyex;
        if (boundary) {
            zin = oz; qxin = oqx; qyin = oqy;
            zex = oz; qxex = oqx; qyex = oqy;
        } else {
            const double nz = win[knb * P + nbk], nqx = win[(3 + knb) * P + nbk], nqy = win[(6 + knb) * P + nbk];
            if (is_in) { zin = oz; qxin = oqx; qyin = oqy; zex = nz; qxex = nqx; qyex = nqy; }
            else { zin = nz; qxin = nqx; qyin = nqy; zex = oz; qxex = oqx; qyex = oqy; }
        }
        if (inf & EI_PRESCRIBED) {                                   // compute2DL.c:149-164
            const int code = (inf >> EI_CODE_SHIFT) & 0xff;
            double qnex;
            prescribed_values(M, ctrl, code, bzeta, bqn, ed, zex, qnex);
            if (M.sentinels && zex < 0) zex = zin;
            if (M.sentinels && qnex > 1000) qnex = qxin * nx + qyin * ny;
            const double qtex = qxin * tx + qyin * ty;
            qxex = (ty * qnex - ny * qtex) * dn;
            qyex = (-tx * qnex + nx * qtex) * dn;
        } else if (boundary) {                                       // :175-191
            zex = oz;
            reflect(qxin, qyin, nx, ny, dn, qxex, qyex);
        }
        double F[3];
        const long long p0 = 2LL * (ed - M.ed_off);   // member-local operand position
        if (WETDRY) {
            const int wd_in = is_in ? wd_self : wd[nbk];
            const int wd_ex = boundary ? wd_self : (is_in ? wd[nbk] : wd_self);
            if (wd_in == 0 && wd_ex == 0) {                          // :203-240
                if (is_in) {
                    double rx, ry;
                    reflect(qxin, qyin, nx, ny, dn, rx, ry);
                    ops.v[2 * k] = roe_flux(F, zin, zin, qxin, rx, qyin, ry, zb, nx, ny, 0.0, g, D);
                    ops.pos[2 * k] = p0;
                    if (boundary) {
                        double Fd[3], sx_, sy_;
                        reflect(qxex, qyex, nx, ny, dn, sx_, sy_);
                        ops.v[2 * k + 1] = roe_flux(Fd, zex, zex, sx_, qxex, sy_, qyex, zb, nx, ny, 0.0, g, D);
                        ops.pos[2 * k + 1] = p0 + 1;
                    }
                } else {
                    double rx, ry;
                    reflect(qxex, qyex, nx, ny, dn, rx, ry);
                    ops.v[2 * k] = roe_flux(F, zex, zex, rx, qxex, ry, qyex, zb, nx, ny, 0.0, g, D);
                    ops.pos[2 * k] = p0 + 1;
                    F[0] = -F[0]; F[1] = -F[1]; F[2] = -F[2];
                }
            } else {                                                 // :241-293
                const int my = is_in ? wd_in : wd_ex;
                ops.v[2 * k] = roe_flux(F, zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, (my == 1) ? g : 0.0, g, D);
                ops.pos[2 * k] = p0;
                // a NaN here comes from the state, so the reference's first check
                // (the g-flux, compute2DL.c:245) would already have stopped
                if ((isnan(F[0]) || isnan(F[1]) || isnan(F[2])) && ed - M.ed_off < err_edge) {
                    err_code = 3;
                    err_edge = ed - M.ed_off;
                }
                if (!is_in) { F[0] = -F[0]; F[1] = -F[1]; F[2] = -F[2]; }
            }
        } else {                                                     // :295-315
            ops.v[2 * k] = roe_flux(F, zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, g, g, D);
            ops.pos[2 * k] = p0;
            if ((isnan(F[0]) || isnan(F[1]) || isnan(F[2])) && ed - M.ed_off < err_edge) {
                err_code = 3;
                err_edge = ed - M.ed_off;
            }
            if (!is_in) { F[0] = -F[0]; F[1] = -F[1]; F[2] = -F[2]; }
        }
        Fh[k] = F[0];
        Fh[3 + k] = F[1];
        Fh[6 + k] = F[2];
    }
}

// max-|lambda| bookkeeping of one element's operands; returns the non-NaN max
__device__ __forceinline__ double lam_fold(const LamOps &ops, Ctrl *ctrl, int slot)
{
    double lam = 0.0;
    long long nanp = -1;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double x = ops.v[i];
        if (ops.pos[i] < 0) continue;
        if (x != x) nanp = (ops.pos[i] > nanp) ? ops.pos[i] : nanp;
        else lam = (x > lam) ? x : lam;
    }
    if (nanp >= 0) atomicMax(&ctrl->nan_pos[slot], nanp);
    return lam;
}

// block max of the thread maxima, one plain store per block into this
// stage's partials (reduced by block_reduce_parts); blockDim.x is 256
__device__ __forceinline__ void lam_commit(double lam, const PhysDev &ph, int slot, int lb)
{
    __shared__ unsigned long long red[4];
    const unsigned long long wm = wave_max((unsigned long long)__double_as_longlong(lam));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = red[0];
#pragma unroll
        for (int i = 1; i < 4; ++i) b = (red[i] > b) ? red[i] : b;
        ph.lamparts[(size_t)slot * ph.maxparts + lb] = b;
    }
}

// Workgroup partial of the reference's in-order max fold over the edge
// kernel's contiguous edge range.  Thread t passes the position of its last
// NaN operand tp (-1: none), the max tm of its operands after it (all of them
// if none; -1.0: none) and its first operand position.  The partial is the
// last NaN position P of the range and the max of the operands after it:
// tm of the thread holding P and of every thread that starts after P.
// blockDim.x is 256.
__device__ __forceinline__ void seg_commit(int tp, double tm, int first, const PhysDev &ph, int slot, int lb)
{
    __shared__ int rp[4];
    __shared__ double rm[4];
    // one pass for P and the max of every tm (the answer when P = -1), a
    // second one only for a workgroup that saw a NaN
    int P = wave_max(tp);
    double M = wave_max(tm);
    if ((threadIdx.x & 63) == 0) { rp[threadIdx.x >> 6] = P; rm[threadIdx.x >> 6] = M; }
    __syncthreads();
    P = rp[0];
    M = rm[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) { P = (rp[i] > P) ? rp[i] : P; M = (rm[i] > M) ? rm[i] : M; }
    if (P >= 0) {   // uniform over the workgroup
        __syncthreads();
        M = wave_max((tp == P || first > P) ? tm : -1.0);
        if ((threadIdx.x & 63) == 0) rm[threadIdx.x >> 6] = M;
        __syncthreads();
        M = rm[0];
#pragma unroll
        for (int i = 1; i < 4; ++i) M = (rm[i] > M) ? rm[i] : M;
    }
    if (threadIdx.x == 0) {
        ph.lamparts[(size_t)slot * ph.maxparts + lb] = (unsigned long long)__double_as_longlong(M);
        ph.nanparts[(size_t)slot * ph.maxparts + lb] = P;
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// the workgroup's largest non-NaN max-|lambda| operand and how many of its
// operands equal it (lane values a, c), for the tie count of a taped step
__device__ __forceinline__ void tie_commit(double a, int c, const PhysDev &ph, int lb)
{
    __shared__ double ra[4];
    __shared__ int rc[4];
    double A = wave_max(a);
    if ((threadIdx.x & 63) == 0) ra[threadIdx.x >> 6] = A;
    __syncthreads();
    A = ra[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) A = (ra[i] > A) ? ra[i] : A;
    const int cc = wave_sum((a == A) ? c : 0);
    if ((threadIdx.x & 63) == 0) rc[threadIdx.x >> 6] = cc;
    __syncthreads();
    if (threadIdx.x == 0) {
        ph.tiemax[lb] = (unsigned long long)__double_as_longlong(A);
        ph.tiecnt[lb] = rc[0] + rc[1] + rc[2] + rc[3];
    }
}

// max of n partials by one block of 256 threads; the result is valid in thread 0
__device__ __forceinline__ unsigned long long block_reduce_parts(const unsigned long long *p, int n)
{
    __shared__ unsigned long long red[4];
    unsigned long long b = 0ULL;
#pragma unroll 8
    for (int i = threadIdx.x; i < n; i += 256) b = (p[i] > b) ? p[i] : b;
    b = wave_max(b);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    b = red[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) b = (red[i] > b) ? red[i] : b;
    return b;
}

// RK update (time_evolution.c:72-88, 176-178, 208-217) and element means
template <int MODE>
__device__ __forceinline__ void rk_update(const MeshDev &m, int e, const double W[9], const double R[9],
                                         double *wout, const double *wbase, double *__restrict__ means,
                                         const Ctrl *ctrl)
{
    const int P = m.pitch;
    const double dt = ctrl->dt;
    double U[9];
    if (MODE == 1) {
#pragma unroll
        for (int f = 0; f < 9; ++f) U[f] = W[f] + R[f] * dt;
    } else {
        const double half = dt / 2;
#pragma unroll
        for (int f = 0; f < 9; ++f) U[f] = (W[f] + wbase[f * P + e]) * 0.5 + R[f] * half;
    }
#pragma unroll
    for (int f = 0; f < 9; ++f) wout[f * P + e] = U[f];
#pragma unroll
    for (int f = 0; f < 3; ++f) means[f * P + e] = (U[3 * f] + U[3 * f + 1] + U[3 * f + 2]) / 3;
}

// ---------------------------------------------------------------------------
// Workgroup -> member.  Element kernels run on pitch/256 workgroups, edge
// kernels on ned_pad/256; every workgroup lies inside one member's range.
// ---------------------------------------------------------------------------
struct ElemLane {
    int lb, mid, e;
    bool valid;
};
__device__ __forceinline__ ElemLane elem_lane_at(const MeshDev &m, int lb)
{
    ElemLane L;
    L.lb = lb;
    L.mid = m.eblk_mem[L.lb];
    L.e = L.lb * 256 + (int)threadIdx.x;
    const MemberDev &M = m.mem[L.mid];
    L.valid = L.e < M.el_off + M.ne;
    return L;
}
__device__ __forceinline__ ElemLane elem_lane(const MeshDev &m) { return elem_lane_at(m, xcd_block()); }
// element kernels whose workgroups carry no partials may run with fewer than
// 256 threads: the element of a thread, its 256-element range and member
__device__ __forceinline__ ElemLane elem_lane_thread(const MeshDev &m)
{
    ElemLane L;
    L.e = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
    L.lb = L.e >> 8;
    L.mid = m.eblk_mem[L.lb];
    const MemberDev &M = m.mem[L.mid];
    L.valid = L.e < M.el_off + M.ne;
    return L;
}

// ---------------------------------------------------------------------------
// k_stage: flux gather + RHS + RK update of one element per thread (fused).
//   MODE 0: RHS only (compute2DL), written to dbg_rhs/dbg_fhat
//   MODE 1: w_out = w_in + RHS*dt                       (time_evolution.c:72-88)
//   MODE 2: w_out = (w_in + w_base)*0.5 + RHS*(dt/2)    (time_evolution.c:176-178,208-217)
//           w_out may alias w_base (each element reads its own slot first).
//   MODE 3: max-|lambda| fold repair after a NaN operand (see Ctrl); returns
//           at once for members whose stage saw no NaN operand.
// ---------------------------------------------------------------------------
template <int MODE, bool WETDRY>
__global__ __launch_bounds__(256) void k_stage(MeshDev m, PhysDev ph, const double *__restrict__ win,
                                               double *wout, const double *wbase,
                                               const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                               const int *__restrict__ wd, double *__restrict__ means,
                                               Ctrl *ctrls, int slot, double *__restrict__ dbg_rhs,
                                               double *__restrict__ dbg_fhat)
{
    const ElemLane L = elem_lane(m);
    const MemberDev &M = m.mem[L.mid];
    Ctrl *ctrl = ctrls + L.mid;
    long long nan_after = -1;
    if (MODE == 3) {
        nan_after = ctrl->nan_pos[slot];
        if (nan_after < 0) return;   // uniform over the workgroup
    }
    const int e = L.e;
    double lam = 0.0;
    bool any = false;
    if (L.valid) {
        const int P = m.pitch;
        double W[9];
#pragma unroll
        for (int f = 0; f < 9; ++f) W[f] = win[f * P + e];
        const int wd_self = wd[e];
        GeoRegs G;
        geometry_of(m, e, ph.g, [&](int i, double v) { G.v[i] = v; });
        double Fh[9];
        LamOps ops;
        int ecode, eedge;
        element_fluxes<WETDRY>(m, ph, M, win, bzeta, bqn, wd, ctrl, e, G, W, wd_self, Fh, ops, ecode, eedge);
        if (MODE == 3) {
#pragma unroll
            for (int i = 0; i < 6; ++i)
                if (ops.pos[i] >= 0 && ops.pos[i] > nan_after) { lam = (ops.v[i] > lam) ? ops.v[i] : lam; any = true; }
        } else {
            lam = lam_fold(ops, ctrl, slot);
            if (ecode) raise_error(ctrl, ecode, eedge, 2 * slot);
            double R[9];
            DivExact D;
            element_rhs(G, W, Fh, wd_self, ph.g, M.rain, ctrl->infil, R, D);
            if (MODE == 0) {
#pragma unroll
                for (int f = 0; f < 9; ++f) { dbg_rhs[f * P + e] = R[f]; dbg_fhat[f * P + e] = Fh[f]; }
            } else {
                rk_update<MODE>(m, e, W, R, wout, wbase, means, ctrl);
            }
        }
    }
    if (MODE == 3) {
        const unsigned long long wm = wave_max((unsigned long long)__double_as_longlong(lam));
        const unsigned long long anyw = __ballot(any);
        if ((threadIdx.x & 63) == 0 && anyw) {
            atomicMax(&ctrl->after_bits[slot], wm);
            atomicOr(&ctrl->any_after[slot], 1);
        }
    } else {
        lam_commit(lam, ph, slot, L.lb);
    }
}

// Split variant, part 1: the three edge fluxes of each element (the edge loop
// of compute2DL in the element-centric orientation) into the Fhat planes.
template <bool WETDRY>
__global__ __launch_bounds__(256) void k_flux(MeshDev m, PhysDev ph, const double *__restrict__ win,
                                              const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                              const int *__restrict__ wd, double *__restrict__ fhat, Ctrl *ctrls,
                                              int slot)
{
    const ElemLane L = elem_lane(m);
    const MemberDev &M = m.mem[L.mid];
    Ctrl *ctrl = ctrls + L.mid;
    const int e = L.e;
    double lam = 0.0;
    if (L.valid) {
        const int P = m.pitch;
        double W[9];
#pragma unroll
        for (int f = 0; f < 9; ++f) W[f] = win[f * P + e];
        const int wd_self = wd[e];
        GeoRegs G;
        geometry_of(m, e, ph.g, [&](int i, double v) { G.v[i] = v; });
        double Fh[9];
        LamOps ops;
        int ecode, eedge;
        element_fluxes<WETDRY>(m, ph, M, win, bzeta, bqn, wd, ctrl, e, G, W, wd_self, Fh, ops, ecode, eedge);
        lam = lam_fold(ops, ctrl, slot);
        if (ecode) raise_error(ctrl, ecode, eedge, 2 * slot);
#pragma unroll
        for (int f = 0; f < 9; ++f) fhat[f * P + e] = Fh[f];
    }
    lam_commit(lam, ph, slot, L.lb);
}

// Canonical geometry of every edge, from el_in's side as the edge loop of
// compute2DL evaluates it (compute2DL.c:98-144): the unit normal (nx, ny),
// the bed at the edge midpoint and 1/(nx*ty - ny*tx), (tx, ty) = (-ny, nx).
// It depends on the nodes only, so it is computed once per mesh (and again
// when the bed changes) instead of from two vertex gathers in every stage.
// The squares and the vertex sum do not depend on the vertex order.
__global__ __launch_bounds__(256) void k_edge_geometry(MeshDev m, double4 *__restrict__ egeo)
{
    const int ed = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (ed >= m.ned_pad) return;
    const int4 E = m.edges[ed];
    const int inf = m.einfo[ed];
    const double4 A = m.nodes[E.z], B = m.nodes[E.w];
    const double len = sqrt((B.x - A.x) * (B.x - A.x) + (B.y - A.y) * (B.y - A.y));
    const double sg = (double)(((inf >> EI_SGN_SHIFT) & 3) - 1);
    const double nx = sg * (B.y - A.y) / len;
    const double ny = sg * (A.x - B.x) / len;
    const double tx = -ny, ty = nx;
    egeo[ed] = make_double4(nx, ny, (A.z + B.z) / 2, 1.0 / (nx * ty - ny * tx));
}

// Edge variant of part 1: one thread per edge, the edge loop of compute2DL
// (compute2DL.c:17-331) as the reference runs it, each Roe flux evaluated
// once and stored for el_in (slot in_k) and, negated, for el_ex (slot ex_k).
// Interior edges need one Roe evaluation, not the two of the element-centric
// gather; both-dry and mixed wet/dry edges evaluate the reference's second
// flux.  The max-|lambda| operands are the edge's own positions 2*edge and
// 2*edge+1, so the fold bookkeeping is that of k_flux.
template <bool WETDRY>
__device__ __forceinline__ void edge_flux_block(const MeshDev &m, const PhysDev &ph, const double *__restrict__ win,
                                                const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                                const int *__restrict__ wd, double *__restrict__ fhat, Ctrl *ctrls,
                                                int slot, int ties, int lb)
{
    const int mid = m.dblk_mem[lb];
    const MemberDev &M = m.mem[mid];
    Ctrl *ctrl = ctrls + mid;
    const int ed = lb * 256 + (int)threadIdx.x;
    int tp = -1, first = -2;
    double tm = -1.0;
    double ta = -1.0;   // largest non-NaN operand of the lane and its count (ties)
    int tc = 0;
    if (ed < M.ed_off + M.ned) {
        const int P = m.pitch;
        const double g = ph.g;
        const int4 E = m.edges[ed];
        const int inf = m.einfo[ed];
        const int el_in = E.x, el_ex = E.y;
        const int kin = inf & ED_KIN_MASK, kex = (inf >> ED_KEX_SHIFT) & 3;
        const bool boundary = (inf & EI_BOUNDARY) != 0;
        const int code = (inf >> EI_CODE_SHIFT) & 0xff;
        const int led = ed - M.ed_off;                 // member-local edge id
        const long long p0 = 2LL * led;
        // the edge's inputs: its geometry, the two midpoint states, boundary data
        struct EdgeIn {
            double4 G;
            double zin, qxin, qyin, zex0, qxex0, qyex0, bz, bq;
            int wd_in, wd_ex;
        };
        auto load = [&](const double4 *egeo, const double *w, const double *bzp, const double *bqp,
                        const int *wdp) {
            EdgeIn in;
            in.G = egeo[ed];
            in.zin = w[kin * P + el_in]; in.qxin = w[(3 + kin) * P + el_in]; in.qyin = w[(6 + kin) * P + el_in];
            in.zex0 = in.zin; in.qxex0 = in.qxin; in.qyex0 = in.qyin;
            if (!boundary) {
                in.zex0 = w[kex * P + el_ex]; in.qxex0 = w[(3 + kex) * P + el_ex]; in.qyex0 = w[(6 + kex) * P + el_ex];
            }
            in.bz = 0.0;
            in.bq = 0.0;
            if (inf & EI_PRESCRIBED) prescribed_values(M, ctrl, code, bzp, bqp, ed, in.bz, in.bq);
            in.wd_in = WETDRY ? wdp[el_in] : 1;
            in.wd_ex = WETDRY ? (boundary ? in.wd_in : wdp[el_ex]) : 1;
            return in;
        };
        LamOps ops;
        double Fi[3], Fx[3];
        bool nan_flux = false;
        // the edge's boundary data and Roe flux(es) with division policy D;
        // its geometry (compute2DL.c:98-144) comes from egeo
        auto eval = [&](const EdgeIn &in, auto &D) {
            const double zin = in.zin, qxin = in.qxin, qyin = in.qyin;
            const int wd_in = in.wd_in, wd_ex = in.wd_ex;
            const double nx = in.G.x, ny = in.G.y;
            const double tx = -ny, ty = nx;
            const double zb = in.G.z;
            const double dn = in.G.w;
            double zex = in.zex0, qxex = in.qxex0, qyex = in.qyex0;
            if (inf & EI_PRESCRIBED) {                                   // compute2DL.c:149-164
                zex = in.bz;
                double qnex = in.bq;
                if (M.sentinels && zex < 0) zex = zin;
                if (M.sentinels && qnex > 1000) qnex = qxin * nx + qyin * ny;
                const double qtex = qxin * tx + qyin * ty;
                qxex = (ty * qnex - ny * qtex) * dn;
                qyex = (-tx * qnex + nx * qtex) * dn;
            } else if (boundary) {                                       // :175-191
                reflect(qxin, qyin, nx, ny, dn, qxex, qyex);
            }
#pragma unroll
            for (int i = 0; i < 6; ++i) { ops.v[i] = 0.0; ops.pos[i] = -1; }
            nan_flux = false;
            if (WETDRY && wd_in == 0 && wd_ex == 0) {                    // :203-240
                double rx, ry, sx, sy;
                reflect(qxin, qyin, nx, ny, dn, rx, ry);
                ops.v[0] = roe_flux(Fi, zin, zin, qxin, rx, qyin, ry, zb, nx, ny, 0.0, g, D);
                ops.pos[0] = p0;
                reflect(qxex, qyex, nx, ny, dn, sx, sy);
                ops.v[1] = roe_flux(Fx, zex, zex, sx, qxex, sy, qyex, zb, nx, ny, 0.0, g, D);
                ops.pos[1] = p0 + 1;
            } else if (WETDRY) {                                         // :241-293
                // the flux of a wet side uses g, of a dry side 0; both give the
                // same |lambda| (the eigenvalues use g, not the flux's gravity)
                const bool in_wet = wd_in == 1, ex_wet = boundary ? in_wet : (wd_ex == 1);
                double Fg[3], F0[3];
                if (in_wet || ex_wet) ops.v[0] = roe_flux(Fg, zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, g, g, D);
                if (!in_wet || !ex_wet) ops.v[0] = roe_flux(F0, zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, 0.0, g, D);
                ops.pos[0] = p0;
#pragma unroll
                for (int j = 0; j < 3; ++j) { Fi[j] = in_wet ? Fg[j] : F0[j]; Fx[j] = ex_wet ? Fg[j] : F0[j]; }
                nan_flux = isnan(Fi[0]) || isnan(Fi[1]) || isnan(Fi[2]) ||
                           (!boundary && (isnan(Fx[0]) || isnan(Fx[1]) || isnan(Fx[2])));
            } else {                                                     // :295-315
                ops.v[0] = roe_flux(Fi, zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, g, g, D);
                ops.pos[0] = p0;
                Fx[0] = Fi[0]; Fx[1] = Fi[1]; Fx[2] = Fi[2];
                nan_flux = isnan(Fi[0]) || isnan(Fi[1]) || isnan(Fi[2]);
            }
        };
        {
            DivFast D;
            eval(load(m.egeo, win, bzeta, bqn, wd), D);
            if (!D.ok()) {
                // an operand outside the fast division's range: plain
                // divisions, on inputs loaded again through pointers the
                // compiler cannot match with the first loads (so none of them
                // stays live through the fast evaluation)
                const double4 *egeo2 = m.egeo;
                const double *win2 = win, *bz2 = bzeta, *bq2 = bqn;
                const int *wd2 = wd;
                asm volatile("" : "+s"(egeo2), "+s"(win2), "+s"(bz2), "+s"(bq2), "+s"(wd2));
                DivExact X;
                eval(load(egeo2, win2, bz2, bq2, wd2), X);
            }
        }
        if (nan_flux) raise_error(ctrl, 3, led, 2 * slot);
        // in-order fold of this edge's operands (positions p0, p0 + 1); a NaN
        // restarts the reference's fold
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (ops.pos[i] < 0) continue;
            const double x = ops.v[i];
            if (x != x) { tp = (int)ops.pos[i]; tm = -1.0; }
            else tm = (x > tm) ? x : tm;
            if (x > ta) { ta = x; tc = 1; }
            else if (x == ta) tc += 1;
        }
        first = (int)p0;
#pragma unroll
        for (int f = 0; f < 3; ++f) fhat[(f * 3 + kin) * P + el_in] = Fi[f];
        if (!boundary) {
#pragma unroll
            for (int f = 0; f < 3; ++f) fhat[(f * 3 + kex) * P + el_ex] = -Fx[f];
        }
    }
    seg_commit(tp, tm, first, ph, slot, lb);
    if (ties) tie_commit(ta, tc, ph, lb);
}

template <bool WETDRY>
__global__ __launch_bounds__(256) void k_edge_flux(MeshDev m, PhysDev ph, const double *__restrict__ win,
                                                   const double *__restrict__ bzeta, const double *__restrict__ bqn,
                                                   const int *__restrict__ wd, double *__restrict__ fhat, Ctrl *ctrls,
                                                   int slot, int ties)
{
    edge_flux_block<WETDRY>(m, ph, win, bzeta, bqn, wd, fhat, ctrls, slot, ties, xcd_block());
}

// Split variants, part 2: inner products, RHS and RK update from the Fhat planes.
//   MODE 0: RHS only into dbg_rhs; MODE 1/2 as in k_stage.
template <int MODE>
__device__ __forceinline__ void rhs_block(const MeshDev &m, const PhysDev &ph, const double *__restrict__ win,
                                          double *wout, const double *wbase, const double *__restrict__ fhat,
                                          const int *__restrict__ wd, double *__restrict__ means, const Ctrl *ctrls,
                                          double *__restrict__ dbg_rhs, const ElemLane &L)
{
    if (!L.valid) return;
    const MemberDev &M = m.mem[L.mid];
    const Ctrl *ctrl = ctrls + L.mid;
    const int e = L.e;
    const int P = m.pitch;
    // the stage input, also kept in this thread's LDS slots for the RK
    // update (so neither registers through the inner products nor a second
    // load from memory hold it)
    __shared__ double wlds[9][256];
    const int tl = (int)threadIdx.x;
    double W[9], Fh[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) {
        W[f] = win[f * P + e];
        wlds[f][tl] = W[f];
    }
#pragma unroll
    for (int f = 0; f < 9; ++f) Fh[f] = fhat[f * P + e];
    const int wd_self = wd[e];
    double R[9];
    {
        DivFast D;
        double4 Pv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) Pv[k] = m.nodes[m.ev[k * P + e]];
        element_rhs_phased(Pv, W, Fh, wd_self, ph.g, M.rain, ctrl->infil, R, D);
        if (!D.ok()) {
            // an operand outside the fast division's range: plain divisions.
            // The inputs are loaded again through pointers the compiler
            // cannot match with the first loads, so nothing of the fast
            // evaluation has to stay live for this path.
            MeshDev m2 = m;
            const double *fh2 = fhat, *w2 = win;
            asm volatile("" : "+s"(m2.nodes), "+s"(m2.ev), "+s"(fh2), "+s"(w2));
            double Fh2[9], W2[9];
#pragma unroll
            for (int f = 0; f < 9; ++f) { Fh2[f] = fh2[f * P + e]; W2[f] = w2[f * P + e]; }
            DivExact X;
            GeoRegs G2;
            geometry_of(m2, e, ph.g, [&](int i, double v) { G2.v[i] = v; }, X);
            element_rhs(G2, W2, Fh2, wd_self, ph.g, M.rain, ctrl->infil, R, X);
        }
    }
    if (MODE == 0) {
#pragma unroll
        for (int f = 0; f < 9; ++f) dbg_rhs[f * P + e] = R[f];
    } else {
        asm volatile("" ::: "memory");
#pragma unroll
        for (int f = 0; f < 9; ++f) W[f] = wlds[f][tl];
        rk_update<MODE>(m, e, W, R, wout, wbase, means, ctrl);
    }
}

template <int MODE>
__global__ __launch_bounds__(256, 2) void k_rhs(MeshDev m, PhysDev ph, const double *__restrict__ win, double *wout,
                                             const double *wbase, const double *__restrict__ fhat,
                                             const int *__restrict__ wd, double *__restrict__ means,
                                             const Ctrl *ctrls, double *__restrict__ dbg_rhs)
{
    rhs_block<MODE>(m, ph, win, wout, wbase, fhat, wd, means, ctrls, dbg_rhs, elem_lane_thread(m));
}

// element means of the current state (standalone limiter calls)
__global__ __launch_bounds__(256) void k_means(MeshDev m, const double *__restrict__ w, double *__restrict__ means)
{
    const ElemLane L = elem_lane(m);
    if (!L.valid) return;
    const int P = m.pitch, e = L.e;
#pragma unroll
    for (int f = 0; f < 3; ++f) means[f * P + e] = (w[(3 * f) * P + e] + w[(3 * f + 1) * P + e] + w[(3 * f + 2) * P + e]) / 3;
}

// node min/max of element means, ascending element order (SlopeLimiter2D.c:106-153).
// ARG: also the first element that attains each bound (-1: the sentinel
// stood), which the limiter's reverse routes the bound adjoints to, as one
// 32-byte record per node: nodearg[8 i + f], f = 0..2 min, 3..5 max, 6..7 pad.
// With lds (a wave's region of NL_CAP x 3 doubles, and NL_CAP ints for ARG)
// the wave first loads the means of the elements its 64 nodes touch (the
// group's sorted list, once each) and each node then reads its entries'
// means there through their packed list positions; the entries, their order
// and the comparisons are those of the CSR loop.  A group whose list is
// longer than NL_CAP, or a null lds, takes the CSR loop over memory.
template <bool ARG>
__device__ __forceinline__ void node_minmax_block(const MeshDev &m, const double *__restrict__ means,
                                                  double *__restrict__ nodemm, int *__restrict__ nodearg, int i,
                                                  double *lds = nullptr)
{
    const int P = m.pitch;
    const int grp = i >> 6, lane = (int)threadIdx.x & 63;
    const bool live = i < m.nn;
    double mn[3] = {99999, 99999, 99999}, mx[3] = {-99999, -99999, -99999};
    int amn[3] = {-1, -1, -1}, amx[3] = {-1, -1, -1};
    bool staged = false;
    if (lds && grp < ((m.nn + 63) >> 6)) {          // uniform over the wave
        const int l0 = m.nl_off[grp], cnt = m.nl_off[grp + 1] - l0;
        if (cnt <= NL_CAP) {
            staged = true;
            // this node's entries and their first two position words (the
            // word array has one word of padding at its end), issued ahead of
            // the wave's loads
            int b = 0, en = 0, w0 = 0;
            unsigned wa = 0, wb = 0;
            if (live) {
                b = m.n2e_off[i]; en = m.n2e_off[i + 1]; w0 = m.loc_off[i];
                wa = m.locw[w0]; wb = m.locw[w0 + 1];
            }
            constexpr int NQ = (NL_CAP + 63) / 64;
            int el[NQ];
#pragma unroll
            for (int r = 0; r < NQ; ++r) el[r] = lane + 64 * r < cnt ? m.nl[l0 + lane + 64 * r] : -1;
            double v[NQ][3];
#pragma unroll
            for (int r = 0; r < NQ; ++r)
#pragma unroll
                for (int f = 0; f < 3; ++f) v[r][f] = el[r] >= 0 ? means[f * P + el[r]] : 0.0;
            int *lid = reinterpret_cast<int *>(lds + 3 * NL_CAP);
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                if (el[r] < 0) continue;
#pragma unroll
                for (int f = 0; f < 3; ++f) lds[f * NL_CAP + lane + 64 * r] = v[r][f];
                if (ARG) lid[lane + 64 * r] = el[r];
            }
            // the wave's stores complete before any lane reads them
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            if (!live) return;
            unsigned word = wa;
            for (int k = 0, t = 0, wi = 1; k < en - b; ++k) {
                if (t == 3) {
                    word = wi == 1 ? wb : m.locw[w0 + wi];
                    ++wi; t = 0;
                }
                const int loc = (int)(word & 1023u);
                word >>= 10; ++t;
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const double a = lds[f * NL_CAP + loc];
                    if (a < mn[f]) { mn[f] = a; if (ARG) amn[f] = lid[loc]; }
                    if (a > mx[f]) { mx[f] = a; if (ARG) amx[f] = lid[loc]; }
                }
            }
        }
    }
    if (!live) return;
    if (!staged) {
        const int b = m.n2e_off[i], en = m.n2e_off[i + 1];
        for (int j = b; j < en; ++j) {
            const int el = m.n2e[j] & N2E_EL_MASK;
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const double a = means[f * P + el];
                if (a < mn[f]) { mn[f] = a; if (ARG) amn[f] = el; }
                if (a > mx[f]) { mx[f] = a; if (ARG) amx[f] = el; }
            }
        }
    }
    const int nn = m.nn;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        nodemm[f * nn + i] = mn[f];
        nodemm[(3 + f) * nn + i] = mx[f];
    }
    if (ARG) {
        int4 *na = reinterpret_cast<int4 *>(nodearg) + 2 * i;
        na[0] = make_int4(amn[0], amn[1], amn[2], amx[0]);
        na[1] = make_int4(amx[1], amx[2], -1, -1);
    }
}

// LDS of the staged node kernels: one region per wave
// (the attaining elements' ids, ARG only, after the means)
template <bool ARG>
constexpr size_t nl_wave_bytes() { return 3 * NL_CAP * sizeof(double) + (ARG ? NL_CAP * sizeof(int) : 0); }

template <bool ARG>
__global__ __launch_bounds__(256) void k_node_minmax(MeshDev m, const double *__restrict__ means,
                                                     double *__restrict__ nodemm, int *__restrict__ nodearg)
{
    extern __shared__ double nl_lds[];
    double *lds = nl_lds + ((int)threadIdx.x >> 6) * (nl_wave_bytes<ARG>() / sizeof(double));
    node_minmax_block<ARG>(m, means, nodemm, nodearg, xcd_block() * (int)blockDim.x + (int)threadIdx.x, lds);
}

// PDop2D on one element (wetDry2D.c:66-256); zeta/qx/qy are the element's
// 3+3+3 midpoint values, modified in place.  Returns an error code or 0.
template <class DV>
__device__ __forceinline__ int pdop_element(double z[3], double qx[3], double qy[3], const double zv[3],
                                            double H0, double VZ, DV &D)
{
    const double bz0 = 0.5 * (zv[0] + zv[1]), bz1 = 0.5 * (zv[1] + zv[2]), bz2 = 0.5 * (zv[2] + zv[0]);
    double mh0 = z[0] + bz0, mh1 = z[1] + bz1, mh2 = z[2] + bz2;
    double Ht[3] = {mh0 - mh1 + mh2, mh0 + mh1 - mh2, -mh0 + mh1 + mh2};
    const double avgH = D.q(Ht[0] + Ht[1] + Ht[2], D.den_const(3.0));
    if (Ht[0] <= H0 || Ht[1] <= H0 || Ht[2] <= H0) {
        if (avgH <= H0) {
            Ht[0] = avgH; Ht[1] = avgH; Ht[2] = avgH;
        } else {
            int mi = 0;
            double mh = Ht[0];
            if (Ht[1] < mh) { mi = 1; mh = Ht[1]; }
            if (Ht[2] < mh) { mi = 2; mh = Ht[2]; }
            int ni = 1;
            double nh = Ht[1];
#pragma unroll
            for (int i = 0; i < 3; ++i) if (Ht[i] < nh && i != mi) { ni = i; nh = Ht[i]; }
            int ri = 2;
#pragma unroll
            for (int i = 2; i >= 0; --i) if (i != mi && i != ni) ri = i;
            // indexed writes in the reference's order
            double h[3] = {Ht[0], Ht[1], Ht[2]};
            h[mi] = H0;
            h[ni] = fmax(H0, nh - 0.5 * (h[mi] - mh));
            h[ri] = h[ri] - (h[mi] - mh) - (h[ni] - nh);
            Ht[0] = h[0]; Ht[1] = h[1]; Ht[2] = h[2];
        }
        if (isnan(Ht[0]) || isnan(Ht[1]) || isnan(Ht[2])) return 6;
        mh0 = 0.5 * (Ht[0] + Ht[1]);
        mh1 = 0.5 * (Ht[1] + Ht[2]);
        mh2 = 0.5 * (Ht[2] + Ht[0]);
        z[0] = mh0 - bz0;
        z[1] = mh1 - bz1;
        z[2] = mh2 - bz2;
    }
    const int f0 = Ht[0] > H0, f1 = Ht[1] > H0, f2 = Ht[2] > H0;
    if (f0 == 0 && f1 == 0 && f2 == 0) {
        qx[0] = 0; qx[1] = 0; qx[2] = 0;
        qy[0] = 0; qy[1] = 0; qy[2] = 0;
    } else if (f0 == 0 || f1 == 0 || f2 == 0) {
        const int f[3] = {f0, f1, f2};
        double qxe[3] = {qx[0] - qx[1] + qx[2], qx[0] + qx[1] - qx[2], -qx[0] + qx[1] + qx[2]};
        double qye[3] = {qy[0] - qy[1] + qy[2], qy[0] + qy[1] - qy[2], -qy[0] + qy[1] + qy[2]};
        double u[3], vv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const typename DV::Den DH = D.den(Ht[i]);
            u[i] = D.q(qxe[i], DH);
            vv[i] = D.q(qye[i], DH);
        }
        int npos = 0;
        double du = 0, dv = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            npos += f[i];
            du += u[i] * (double)(1 - f[i]) * (double)(fabs(qxe[i]) > VZ);
            dv += vv[i] * (double)(1 - f[i]) * (double)(fabs(qye[i]) > VZ);
        }
        int err = 0;
        const typename DV::Den Dn = D.den((double)npos);
        const double dun = D.q(du, Dn), dvn = D.q(dv, Dn);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            u[i] = (double)f[i] * (u[i] + dun);
            vv[i] = (double)f[i] * (vv[i] + dvn);
            if (isnan(u[i]) || isnan(vv[i])) err = 7;
        }
        if (err) return err;
#pragma unroll
        for (int i = 0; i < 3; ++i) { qxe[i] = Ht[i] * u[i]; qye[i] = Ht[i] * vv[i]; }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int i2 = (i + 1) % 3;
            qx[i] = 0.5 * (qxe[i] + qxe[i2]);
            qy[i] = 0.5 * (qye[i] + qye[i2]);
        }
    }
    return 0;
}

// wetDryStatus2D (wetDry2D.c:16-60) of one element from its zeta midpoints
// (a, b, c), vertex depths zz and previous flag
template <class DV>
__device__ __forceinline__ int wd_status_of(double a, double b, double c, const double zz[3], double H0, int old,
                                            DV &D)
{
    const double ze[3] = {a - b + c, a + b - c, -a + b + c};
    const double h[3] = {ze[0] + zz[0], ze[1] + zz[1], ze[2] + zz[2]};
    const double avgH = D.q(h[0] + h[1] + h[2], D.den_const(3.0));
    int mi = 0;
    double mh = h[0];
    if (h[1] > mh) { mi = 1; mh = h[1]; }
    if (h[2] > mh) { mi = 2; mh = h[2]; }
    const double zmax = ze[mi];
    double zmin = zz[0];
    if (zz[1] < zmin) zmin = zz[1];
    if (zz[2] < zmin) zmin = zz[2];
    const int t1 = (avgH - H0) > 0;
    const int t2 = (zmax - (H0 - zmin)) > 0;
    return (old == -1) ? t1 * t2 : old * t1 + (1 - old) * t1 * t2;
}

// limiter element pass (SlopeLimiter2D.c:157-343), optionally followed by PDop2D
// win -> wout (may alias).  With PDOP, the wet/dry status of the limited
// state follows (the reference's next wetDryStatus2D call, which always comes
// right after the limiter): wd_out from wd_in (may alias).
template <bool PDOP>
__device__ __forceinline__ void limit_block(const MeshDev &m, const PhysDev &ph, const double *win, double *wout,
                                            const double *__restrict__ nodemm,
                                            unsigned char *__restrict__ modify_out, Ctrl *ctrls, int stage,
                                            const int *wd_in, int *wd_out, const ElemLane &L, int step_adj)
{
    if (!L.valid) return;
    const int e = L.e;
    const int P = m.pitch, nn = m.nn;
    const double H0 = ph.h0;
    struct LimIn {
        double zv[3], S[9], Vmin[3][3], Vmax[3][3];
        int old;
    };
    auto load = [&](const int *ev, const double *zn, const double *w, const double *nm, const int *wdi) {
        LimIn in;
        int nd[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) { nd[k] = ev[k * P + e]; in.zv[k] = zn[nd[k]]; }
#pragma unroll
        for (int f = 0; f < 9; ++f) in.S[f] = w[f * P + e];
#pragma unroll
        for (int iv = 0; iv < 3; ++iv)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                in.Vmin[iv][k] = nm[iv * nn + nd[k]];
                in.Vmax[iv][k] = nm[(3 + iv) * nn + nd[k]];
            }
        in.old = (PDOP && wd_out) ? wdi[e] : 0;
        return in;
    };
    double S[9];
    int modify = 0, err = 0, wdv = 0;
    // the limiter pass, PDop and the wet/dry status with division policy D
    auto eval = [&](const LimIn &in, auto &D) {
        const double *zv = in.zv;
#pragma unroll
        for (int f = 0; f < 9; ++f) S[f] = in.S[f];
        modify = 0;
        const auto D3 = D.den_const(3.0);
#pragma unroll
        for (int iv = 0; iv < 3; ++iv) {
            const double w1 = S[3 * iv], w2 = S[3 * iv + 1], w3 = S[3 * iv + 2];
            double V[3] = {w1 - w2 + w3, w1 + w2 - w3, -w1 + w2 + w3};
            const double *Vmin = in.Vmin[iv], *Vmax = in.Vmax[iv];
            const double avg = D.q(w1 + w2 + w3, D3);
            if (iv == 0) {
                const double Ha = V[0] + zv[0], Hb = V[1] + zv[1], Hc = V[2] + zv[2];
                const double Hm0 = Vmin[0] + zv[0], Hm1 = Vmin[1] + zv[1], Hm2 = Vmin[2] + zv[2];
                if (Ha > H0 && Hb > H0 && Hc > H0 && Hm0 > H0 && Hm1 > H0 && Hm2 > H0) modify = 1;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) V[k] = fmax(DGR_MIN(V[k], Vmax[k]), Vmin[k]);
#pragma unroll
            for (int ll = 0; ll < 3; ++ll) {
                const double sumloc = D.q(V[0] + V[1] + V[2], D3);
                double sumdiff = (sumloc - avg) * 3;
                const int sd = isign(sumdiff);
                const double diff[3] = {(V[0] - avg) * sd, (V[1] - avg) * sd, (V[2] - avg) * sd};
                double kdp = (double)((diff[0] > 0) + (diff[1] > 0) + (diff[2] > 0));
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    // sumdiff*sd / max(1, kdp) with kdp in {0,1,2,3}: x/1 = x and
                    // x/2 = x*0.5 exactly.  kdp = 3 means all three diffs are
                    // positive, and k = 0 takes it down to 2: only k = 0 can
                    // divide by 3.
                    double redfac;
                    if (diff[k] > 0) {
                        const double num = sumdiff * sd;
                        redfac = (k == 0 && kdp > 2.0) ? D.q(num, D3) : (kdp > 1.0) ? num * 0.5 : num;
                        kdp -= 1;
                    } else redfac = 0;
                    const double redmax = (sd > 0) ? V[k] - Vmin[k] : Vmax[k] - V[k];
                    redfac = DGR_MIN(redfac, redmax);
                    sumdiff -= redfac * sd;
                    V[k] -= redfac * sd;
                }
            }
            const double mid1 = (V[0] + V[1]) / 2, mid2 = (V[1] + V[2]) / 2, mid3 = (V[2] + V[0]) / 2;
            if (iv == 0) {
                const double Ha = V[0] + zv[0], Hb = V[1] + zv[1], Hc = V[2] + zv[2];
                if (Ha < H0 || Hb < H0 || Hc < H0) modify = 0;
            }
            if (modify) { S[3 * iv] = mid1; S[3 * iv + 1] = mid2; S[3 * iv + 2] = mid3; }
        }
        err = 0;
        if (PDOP) {
            err = pdop_element(&S[0], &S[3], &S[6], zv, H0, ph.velzero, D);
            if (wd_out) wdv = wd_status_of(S[0], S[1], S[2], zv, H0, in.old, D);
        }
    };
    {
        DivFast D;
        eval(load(m.ev, m.zn, win, nodemm, wd_in), D);
        if (!D.ok()) {
            // an operand outside the fast division's range: plain divisions on
            // inputs loaded again (nothing of the fast pass stays live for it)
            const int *ev2 = m.ev, *wd2 = wd_in;
            const double *zn2 = m.zn;
            const double *win2 = win, *nm2 = nodemm;
            asm volatile("" : "+s"(ev2), "+s"(zn2), "+s"(win2), "+s"(nm2), "+s"(wd2));
            DivExact X;
            eval(load(ev2, zn2, win2, nm2, wd2), X);
        }
    }
    if (err) raise_error(ctrls + L.mid, err, e - m.mem[L.mid].el_off, 2 * stage + 1, step_adj);
    if (PDOP && wd_out) wd_out[e] = wdv;
#pragma unroll
    for (int f = 0; f < 9; ++f) wout[f * P + e] = S[f];
    if (modify_out) modify_out[e] = (unsigned char)modify;
}

template <bool PDOP>
__global__ __launch_bounds__(256) void k_limit(MeshDev m, PhysDev ph, const double *win, double *wout,
                                               const double *__restrict__ nodemm,
                                               unsigned char *__restrict__ modify_out, Ctrl *ctrls, int stage,
                                               const int *wd_in, int *wd_out, int step_adj)
{
    limit_block<PDOP>(m, ph, win, wout, nodemm, modify_out, ctrls, stage, wd_in, wd_out, elem_lane_thread(m),
                      step_adj);
}

// standalone PDop2D
__global__ __launch_bounds__(256) void k_pdop(MeshDev m, PhysDev ph, double *w, Ctrl *ctrls)
{
    const ElemLane L = elem_lane(m);
    if (!L.valid) return;
    const int e = L.e;
    const int P = m.pitch;
    double zv[3], S[9];
#pragma unroll
    for (int k = 0; k < 3; ++k) zv[k] = m.zn[m.ev[k * P + e]];
#pragma unroll
    for (int f = 0; f < 9; ++f) S[f] = w[f * P + e];
    DivExact D;
    const int err = pdop_element(&S[0], &S[3], &S[6], zv, ph.h0, ph.velzero, D);
    if (err) raise_error(ctrls + L.mid, err, e - m.mem[L.mid].el_off, 1);
#pragma unroll
    for (int f = 0; f < 9; ++f) w[f * P + e] = S[f];
}

// standalone wetDryStatus2D: wd_out from the state and wd_in (may alias)
__global__ __launch_bounds__(256) void k_wd_status(MeshDev m, PhysDev ph, const double *__restrict__ w,
                                                   const int *wd_in, int *wd_out)
{
    const ElemLane L = elem_lane(m);
    if (!L.valid) return;
    const int e = L.e;
    const int P = m.pitch;
    double zz[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) zz[k] = m.zn[m.ev[k * P + e]];
    DivExact D;
    wd_out[e] = wd_status_of(w[e], w[P + e], w[2 * P + e], zz, ph.h0, wd_in[e], D);
}

__device__ __forceinline__ void clear_fold(Ctrl *c)
{
    c->maxlam[0] = 0ULL;
    c->maxlam[1] = 0ULL;
    c->nan_pos[0] = -1;
    c->nan_pos[1] = -1;
    c->after_bits[0] = 0ULL;
    c->after_bits[1] = 0ULL;
    c->any_after[0] = 0;
    c->any_after[1] = 0;
}

// per-step bookkeeping (time_evolution.c:311-402), one thread per member:
// begin evaluates the tide at the step's start time; end advances t and sets
// the next dt from stage 2's max |lambda|.
// the NaN bookkeeping of the element-kernel folds (Ctrl::maxlam is written by
// the reductions before it is read)
__device__ __forceinline__ void clear_nan_fold(Ctrl *c)
{
    c->nan_pos[0] = -1;
    c->nan_pos[1] = -1;
    c->after_bits[0] = 0ULL;
    c->after_bits[1] = 0ULL;
    c->any_after[0] = 0;
    c->any_after[1] = 0;
}

// the time-dependent boundary values of the step starting at c->t: the tide
// and the hydrographs (boundary_conditions.c:10-83 is called once per step)
__device__ __forceinline__ void step_tide(const MeshDev &m, const MemberDev &M, Ctrl *c)
{
    if (M.tide_code)
        c->tide = M.tide_datum + M.tide_amp * dgr_sin(6.283185307179586 * c->t / M.tide_period + M.tide_phase);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int n = M.hyd_n[k];
        if (n) {
            const double *h = m.hydro + M.hyd_off[k];
            c->hyd_z[k] = hydro_eval(n, h, h + n, c->t);
            c->hyd_q[k] = hydro_eval(n, h, h + 2 * n, c->t);
        }
    }
    // the step's mean infiltration rate (F(t + dt) - F(t)) / dt
    c->infil = (M.infil_K > 0.0 && c->dt > 0.0) ? (ga_cumulative(M.infil_K, M.infil_P, c->t + c->dt) -
                                                   ga_cumulative(M.infil_K, M.infil_P, c->t)) / c->dt
                                                : 0.0;
}

__global__ __launch_bounds__(64) void k_step_begin(MeshDev m, Ctrl *ctrls)
{
    const int mid = blockIdx.x * 64 + threadIdx.x;
    if (mid >= m.nmem) return;
    const MemberDev &M = m.mem[mid];
    Ctrl *c = ctrls + mid;
    step_tide(m, M, c);
    clear_fold(c);
    c->step += 1;
}

// oneTimeStep2D with a host dt: no tide update, only the step bookkeeping
__global__ __launch_bounds__(64) void k_step_mark(MeshDev m, Ctrl *ctrls, double dt)
{
    const int mid = blockIdx.x * 64 + threadIdx.x;
    if (mid >= m.nmem) return;
    Ctrl *c = ctrls + mid;
    c->dt = dt;
    clear_fold(c);
    c->step += 1;
}

// the partials one member's workgroups wrote for a slot
__device__ __forceinline__ unsigned long long member_lam(const MeshDev &m, const PhysDev &ph, int mid, int slot,
                                                          int kind)
{
    const MemberDev &M = m.mem[mid];
    const int b0 = kind == PARTS_EDGE ? M.dblk0 : M.eblk0;
    const int nb = kind == PARTS_EDGE ? M.dblks : M.eblks;
    return block_reduce_parts(ph.lamparts + (size_t)slot * ph.maxparts + b0, nb);
}

// The reference's in-order max fold (compute2DL.c:232-238,289-292) over one
// member's operands from the edge kernel's workgroup partials (seg_commit);
// member-local workgroup b holds operand positions [512 b, 512 b + 512).  The
// fold starts at 0; a NaN operand restarts it, so the result is the max of
// the operands after the last NaN, or NaN if none follows.  One workgroup of
// 256 threads; the result is valid in every thread.
// With tie != nullptr (a taped step of the edge kernel) the same pass also
// reduces the tie partials (tie_commit): the largest non-NaN operand A and
// how many operands equal it, as (max, count) pairs (a larger max takes its
// count, an equal one adds it; exact, so any order gives the same pair);
// *tie gets the count when A is the fold's result, else -1 (a NaN operand
// restarted the fold).
__device__ __forceinline__ double member_fold_edges(const MeshDev &m, const PhysDev &ph, int mid, int slot,
                                                    int *tie = nullptr)
{
    __shared__ int rp[4], rc[4];
    __shared__ double rm[4], ra[4];
    const MemberDev &M = m.mem[mid];
    const int *np = ph.nanparts + (size_t)slot * ph.maxparts + M.dblk0;
    const unsigned long long *lp = ph.lamparts + (size_t)slot * ph.maxparts + M.dblk0;
    const unsigned long long *tp = ph.tiemax + M.dblk0;
    const int *tn = ph.tiecnt + M.dblk0;
    // one pass for P and the max of all partials (the answer when P = -1),
    // a second one only when the member saw a NaN operand
    int P = -1, C = 0;
    double R = -1.0, A = -1.0;
    // unrolled so that several partials' loads are in flight at once (one
    // workgroup folds every partial of the member)
    if (tie) {
#pragma unroll 8
        for (int b = threadIdx.x; b < M.dblks; b += 256) {
            const double v = __longlong_as_double((long long)lp[b]);
            P = (np[b] > P) ? np[b] : P;
            R = (v > R) ? v : R;
            const double a = __longlong_as_double((long long)tp[b]);
            const int c = tn[b];
            C = (a > A) ? c : (a == A) ? C + c : C;
            A = (a > A) ? a : A;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double a = __shfl_xor(A, o, 64);
            const int c = __shfl_xor(C, o, 64);
            C = (a > A) ? c : (a == A) ? C + c : C;
            A = (a > A) ? a : A;
        }
        if ((threadIdx.x & 63) == 0) { ra[threadIdx.x >> 6] = A; rc[threadIdx.x >> 6] = C; }
    } else {
#pragma unroll 8
        for (int b = threadIdx.x; b < M.dblks; b += 256) {
            const double v = __longlong_as_double((long long)lp[b]);
            P = (np[b] > P) ? np[b] : P;
            R = (v > R) ? v : R;
        }
    }
    P = wave_max(P);
    R = wave_max(R);
    if ((threadIdx.x & 63) == 0) { rp[threadIdx.x >> 6] = P; rm[threadIdx.x >> 6] = R; }
    __syncthreads();
    P = rp[0];
    R = rm[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) { P = (rp[i] > P) ? rp[i] : P; R = (rm[i] > R) ? rm[i] : R; }
    if (tie) {
        A = ra[0];
        C = rc[0];
#pragma unroll
        for (int i = 1; i < 4; ++i) {
            C = (ra[i] > A) ? rc[i] : (ra[i] == A) ? C + rc[i] : C;
            A = (ra[i] > A) ? ra[i] : A;
        }
    }
    if (P < 0) {
        const double ml = (R > 0.0) ? R : 0.0;
        if (tie) *tie = (A == ml) ? C : -1;
        return ml;
    }
    __syncthreads();
    R = -1.0;
#pragma unroll 8
    for (int b = threadIdx.x; b < M.dblks; b += 256) {
        const double v = __longlong_as_double((long long)lp[b]);
        if ((np[b] == P || 512LL * b > P) && v > R) R = v;
    }
    R = wave_max(R);
    if ((threadIdx.x & 63) == 0) rm[threadIdx.x >> 6] = R;
    __syncthreads();
    R = rm[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) R = (rm[i] > R) ? rm[i] : R;
    const double ml = (R < 0.0) ? __longlong_as_double(0x7ff8000000000000LL) : R;
    if (tie) *tie = (A == ml) ? C : -1;
    return ml;
}

// the fold of one member and stage from the last flux launch's partials,
// stored resolved in Ctrl::maxlam[slot] (thread 0)
__device__ __forceinline__ double resolve_lam(const MeshDev &m, const PhysDev &ph, Ctrl *c, int mid, int slot,
                                              int kind, int *tie = nullptr)
{
    if (kind == PARTS_EDGE) {
        const double ml = member_fold_edges(m, ph, mid, slot, tie);
        if (threadIdx.x == 0) {
            c->maxlam[slot] = (unsigned long long)__double_as_longlong(ml);
            c->nan_pos[slot] = -1;
        }
        return ml;
    }
    const unsigned long long b = member_lam(m, ph, mid, slot, kind);
    if (threadIdx.x != 0) return 0.0;
    c->maxlam[slot] = b;
    const double ml = maxlam_of(*c, slot);
    c->maxlam[slot] = (unsigned long long)__double_as_longlong(ml);
    c->nan_pos[slot] = -1;
    return ml;
}

// the member's count of stage-2 max-|lambda| operands equal to ml (the
// resolved fold) from the taped edge launch's tie partials, or -1 when the
// fold saw a NaN operand (the largest non-NaN operand is then not ml).  All
// threads of the workgroup call it; the result is valid in thread 0.
__device__ __forceinline__ int member_ties(const MeshDev &m, const PhysDev &ph, int mid, double ml)
{
    __shared__ double ra[4];
    __shared__ int rc[4];
    const MemberDev &M = m.mem[mid];
    const unsigned long long *tp = ph.tiemax + M.dblk0;
    const int *tn = ph.tiecnt + M.dblk0;
    double A = -1.0;
#pragma unroll 8
    for (int b = threadIdx.x; b < M.dblks; b += 256) {
        const double v = __longlong_as_double((long long)tp[b]);
        A = (v > A) ? v : A;
    }
    A = wave_max(A);
    if ((threadIdx.x & 63) == 0) ra[threadIdx.x >> 6] = A;
    __syncthreads();
    A = ra[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) A = (ra[i] > A) ? ra[i] : A;
    int C = 0;
#pragma unroll 8
    for (int b = threadIdx.x; b < M.dblks; b += 256)
        if (__longlong_as_double((long long)tp[b]) == A) C += tn[b];
    C = wave_sum(C);
    if ((threadIdx.x & 63) == 0) rc[threadIdx.x >> 6] = C;
    __syncthreads();
    C = rc[0] + rc[1] + rc[2] + rc[3];
    return (A == ml) ? C : -1;
}

// one workgroup per member.  With next != 0 the next step's bookkeeping
// follows (the tide at the new time, the NaN bookkeeping cleared, the step
// counted), which k_step_begin does otherwise.  Optional step records: the
// step's max |lambda| into rec_end, the next step's t, dt and tide into
// rec_next (after the next step's bookkeeping).
__device__ __forceinline__ void step_end_member(const MeshDev &m, const PhysDev &ph, Ctrl *ctrls, int kind, int next,
                                                StepRec *rec_end, StepRec *rec_next, int ties, int mid)
{
    Ctrl *c = ctrls + mid;
    // taped steps of the edge kernel count the max's ties in the fold's pass
    int ntie = 0;
    const bool fused = ties && kind == PARTS_EDGE;
    const double ml = resolve_lam(m, ph, c, mid, 1, kind, fused ? &ntie : nullptr);
    if (ties && !fused) ntie = member_ties(m, ph, mid, ml);
    if (threadIdx.x != 0) return;
    if (ties && rec_end) rec_end[mid].ntie = ntie;
    const MemberDev &M = m.mem[mid];
    const double time = c->t + c->dt;
    double h = M.cfl * M.min_edge / ml;
    if ((time + h) > M.final_time) h = M.final_time - time;
    c->t = time;
    c->dt = h;
    if (next) {
        step_tide(m, M, c);
        clear_nan_fold(c);
        c->step += 1;
    }
    if (rec_end) rec_end[mid].lam2 = c->maxlam[1];
    if (rec_next) {
        rec_next[mid].t = c->t;
        rec_next[mid].dt = c->dt;
        rec_next[mid].tide = c->tide;
    }
}

__global__ __launch_bounds__(256) void k_step_end(MeshDev m, PhysDev ph, Ctrl *ctrls, int kind, int next,
                                                  StepRec *rec_end, StepRec *rec_next, int ties)
{
    step_end_member(m, ph, ctrls, kind, next, rec_end, rec_next, ties, (int)blockIdx.x);
}

// The stage-2 node bounds of a run step with the step end folded in (one
// launch fewer per step): the first nf workgroups (nmem rounded up to a
// multiple of 8, so that the node workgroups keep their XCD mapping) end
// one member's step each, as k_step_end with next = 1 does (with the step
// records and tie counts of a taped step); the others compute node bounds
// (ARG: and the elements attaining them).  Every input of the step end (the
// stage-2 flux partials) is complete when this launch starts, and of what
// it writes the following stage-2 limiter reads only the step count (its
// error keys subtract the step it has advanced).
template <bool ARG>
__global__ __launch_bounds__(256) void k_node_minmax_end(MeshDev m, PhysDev ph, const double *__restrict__ means,
                                                         double *__restrict__ nodemm, int *__restrict__ nodearg,
                                                         Ctrl *ctrls, int kind, int nf, StepRec *rec_end,
                                                         StepRec *rec_next, int ties)
{
    const int b = (int)blockIdx.x;
    if (b < nf) {
        if (b < m.nmem) step_end_member(m, ph, ctrls, kind, 1, rec_end, rec_next, ties, b);
        return;
    }
    extern __shared__ double nl_lds[];
    double *lds = nl_lds + ((int)threadIdx.x >> 6) * (nl_wave_bytes<ARG>() / sizeof(double));
    node_minmax_block<ARG>(m, means, nodemm, nodearg, xcd_map(b - nf, (int)gridDim.x - nf) * 256 + (int)threadIdx.x,
                           lds);
}

// ---------------------------------------------------------------------------
// Persistent variant: one cooperative launch runs a whole RK stage of every
// member: edge fluxes, grid barrier, RHS + RK update, barrier, node bounds,
// barrier, limiter (+ PDop and the wet/dry status), and in stage 2 a barrier
// and the step end.  The grid is as many workgroups as fit on the device at
// once (a multiple of 8); workgroup b runs virtual workgroups b, b + G, ...
// of each phase, which lie on b's XCD, through the same XCD mapping as the
// separate launches.  Same device functions, same bits.
// ---------------------------------------------------------------------------
template <int MODE, bool WETDRY>
__global__ __launch_bounds__(256, 2) void k_stage_persist(MeshDev m, PhysDev ph, PersistArgs a)
{
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int G = (int)gridDim.x, b = (int)blockIdx.x;
    const int nv_ed = m.ned_pad / 256, nv_el = m.pitch / 256, nv_nd = (m.nn + 255) / 256;
    const int slot = MODE - 1;
    for (int v = b; v < nv_ed; v += G) {
        edge_flux_block<WETDRY>(m, ph, a.win, a.bzeta, a.bqn, a.wd, a.fhat, a.ctrls, slot, 0, xcd_map(v, nv_ed));
        __syncthreads();   // the workgroup reductions' LDS is reused by the next virtual workgroup
    }
    grid.sync();
    for (int v = b; v < nv_el; v += G)
        rhs_block<MODE>(m, ph, a.win, a.u, a.wbase, a.fhat, a.wd, a.means, a.ctrls, nullptr,
                        elem_lane_at(m, xcd_map(v, nv_el)));
    grid.sync();
    for (int v = b; v < nv_nd; v += G)
        node_minmax_block<false>(m, a.means, a.nodemm, nullptr, xcd_map(v, nv_nd) * 256 + (int)threadIdx.x);
    grid.sync();
    for (int v = b; v < nv_el; v += G)
        limit_block<WETDRY>(m, ph, a.u, a.lout, a.nodemm, a.modify, a.ctrls, slot, a.wd_lin, a.wd_lout,
                            elem_lane_at(m, xcd_map(v, nv_el)), 0);
    if (MODE == 2) {
        grid.sync();
        for (int mid = b; mid < m.nmem; mid += G) {
            step_end_member(m, ph, a.ctrls, PARTS_EDGE, a.next, nullptr, nullptr, 0, mid);
            __syncthreads();
        }
    }
}

// Ctrl::maxlam[slot] of every member from the partials of the last flux launch
__global__ __launch_bounds__(256) void k_lam_reduce(MeshDev m, PhysDev ph, Ctrl *ctrls, int slot, int kind)
{
    const int mid = blockIdx.x;
    (void)resolve_lam(m, ph, ctrls + mid, mid, slot, kind);
}

__global__ __launch_bounds__(64) void k_reset_maxlam(MeshDev m, Ctrl *ctrls)
{
    const int mid = blockIdx.x * 64 + threadIdx.x;
    if (mid < m.nmem) clear_fold(ctrls + mid);
}

constexpr int NSELF = 9;   // outputs per input pair
__global__ void k_math_selftest(int n, const double *x, const double *y, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = x[i], b = y[i];
    out[NSELF * i + 0] = sqrt(a);
    out[NSELF * i + 1] = a / b;
    out[NSELF * i + 2] = dgr_pow43(a);
    out[NSELF * i + 3] = dgr_sin(a);
    out[NSELF * i + 4] = fma(a, b, a);
    DivFast D;
    const double q1 = D.q(a, D.den(b));
    const bool ok1 = D.ok();
    DivFast D3;
    const double q3 = D3.q(a, D3.den_const(3.0));
    const bool ok3 = D3.ok();
    // the fast quotient where its range check passes, else plain division
    out[NSELF * i + 5] = ok1 ? q1 : a / b;
    out[NSELF * i + 6] = ok3 ? q3 : a / 3.0;
    out[NSELF * i + 7] = dgr_log1p(fabs(a));
    out[NSELF * i + 8] = ga_cumulative(1e-6 + 1e-5 * fabs(b), 0.033, fabs(a));
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 grid_for(int n) { return dim3((unsigned)((n + 255) / 256)); }
static inline dim3 member_grid(const MeshDev &m) { return dim3((unsigned)((m.nmem + 63) / 64)); }

// variant 0: fused k_stage; 1: k_flux (element-centric) + k_rhs;
// 2: k_edge_flux (edge-centric) + k_rhs.  MODE 3 always runs k_stage.
// Returns the workgroup space the launch wrote its max-|lambda| partials in.
int launch_stage(int variant, int mode, bool wetdry, const MeshDev &m, const PhysDev &ph, const double *win,
                 double *wout, const double *wbase, const double *bzeta, const double *bqn, const int *wd,
                 double *means, double *fhat, Ctrl *ctrl, int slot, double *dbg_rhs, double *dbg_fhat,
                 hipStream_t s, bool ties)
{
    const dim3 g = grid_for(m.pitch), b(256);
    const dim3 gw((unsigned)(m.pitch / m.wg)), bw((unsigned)m.wg);
#define DGR_LS(MD, WDF) \
    hipLaunchKernelGGL((k_stage<MD, WDF>), g, b, 0, s, m, ph, win, wout, wbase, bzeta, bqn, wd, means, ctrl, slot, dbg_rhs, dbg_fhat)
#define DGR_LR(MD) \
    hipLaunchKernelGGL((k_rhs<MD>), gw, bw, 0, s, m, ph, win, wout, wbase, mode == 0 ? dbg_fhat : fhat, wd, means, ctrl, dbg_rhs)
    if (variant == DGR_VARIANT_PERSIST) variant = 2;   // single stages: the separate launches
    if (variant == 0 || mode == 3) {
        if (wetdry) {
            if (mode == 0) DGR_LS(0, true); else if (mode == 1) DGR_LS(1, true);
            else if (mode == 2) DGR_LS(2, true); else DGR_LS(3, true);
        } else {
            if (mode == 0) DGR_LS(0, false); else if (mode == 1) DGR_LS(1, false);
            else if (mode == 2) DGR_LS(2, false); else DGR_LS(3, false);
        }
        return PARTS_ELEM;
    }
    double *fh = mode == 0 ? dbg_fhat : fhat;
    int kind;
    if (variant == 1) {
        if (wetdry) hipLaunchKernelGGL(k_flux<true>, g, b, 0, s, m, ph, win, bzeta, bqn, wd, fh, ctrl, slot);
        else hipLaunchKernelGGL(k_flux<false>, g, b, 0, s, m, ph, win, bzeta, bqn, wd, fh, ctrl, slot);
        kind = PARTS_ELEM;
    } else {
        const dim3 ge = grid_for(m.ned_pad);
        const int t = ties ? 1 : 0;
        if (wetdry) hipLaunchKernelGGL(k_edge_flux<true>, ge, b, 0, s, m, ph, win, bzeta, bqn, wd, fh, ctrl, slot, t);
        else hipLaunchKernelGGL(k_edge_flux<false>, ge, b, 0, s, m, ph, win, bzeta, bqn, wd, fh, ctrl, slot, t);
        kind = PARTS_EDGE;
    }
    if (mode == 0) DGR_LR(0); else if (mode == 1) DGR_LR(1); else DGR_LR(2);
#undef DGR_LS
#undef DGR_LR
    return kind;
}

void launch_means(const MeshDev &m, const double *w, double *means, hipStream_t s)
{
    hipLaunchKernelGGL(k_means, grid_for(m.pitch), dim3(256), 0, s, m, w, means);
}

void launch_edge_geometry(const MeshDev &m, double4 *egeo, hipStream_t s)
{
    hipLaunchKernelGGL(k_edge_geometry, grid_for(m.ned_pad), dim3(256), 0, s, m, egeo);
}

void launch_node_minmax(const MeshDev &m, const double *means, double *nodemm, int *nodearg, hipStream_t s)
{
    const dim3 g((unsigned)((m.nn + m.wg - 1) / m.wg)), b((unsigned)m.wg);
    const size_t w = (size_t)(m.wg / 64);
    if (nodearg) hipLaunchKernelGGL(k_node_minmax<true>, g, b, w * nl_wave_bytes<true>(), s, m, means, nodemm, nodearg);
    else hipLaunchKernelGGL(k_node_minmax<false>, g, b, w * nl_wave_bytes<false>(), s, m, means, nodemm, nodearg);
}

void launch_limit(bool pdop, const MeshDev &m, const PhysDev &ph, const double *win, double *wout,
                  const double *nodemm, unsigned char *modify, Ctrl *ctrl, int stage, const int *wd_in, int *wd_out,
                  hipStream_t s, int step_adj)
{
    const dim3 g((unsigned)(m.pitch / m.wg)), b((unsigned)m.wg);
    if (pdop)
        hipLaunchKernelGGL(k_limit<true>, g, b, 0, s, m, ph, win, wout, nodemm, modify, ctrl, stage, wd_in, wd_out,
                           step_adj);
    else
        hipLaunchKernelGGL(k_limit<false>, g, b, 0, s, m, ph, win, wout, nodemm, modify, ctrl, stage, wd_in, wd_out,
                           step_adj);
}

void launch_pdop(const MeshDev &m, const PhysDev &ph, double *w, Ctrl *ctrl, hipStream_t s)
{
    hipLaunchKernelGGL(k_pdop, grid_for(m.pitch), dim3(256), 0, s, m, ph, w, ctrl);
}

void launch_wd_status(const MeshDev &m, const PhysDev &ph, const double *w, const int *wd_in, int *wd_out,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_wd_status, grid_for(m.pitch), dim3(256), 0, s, m, ph, w, wd_in, wd_out);
}

void launch_step_begin(const MeshDev &m, Ctrl *c, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_begin, member_grid(m), dim3(64), 0, s, m, c);
}

// workgroups of a persistent launch: as many as fit at once, a multiple of 8
static int persist_grid(const void *kernel)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess) return 0;
    return (per_cu * cus) & ~7;
}

hipError_t launch_stage_persist(int mode, bool wetdry, const MeshDev &m, const PhysDev &ph, const PersistArgs &a,
                                hipStream_t s)
{
    const void *k = mode == 1 ? (wetdry ? (const void *)k_stage_persist<1, true> : (const void *)k_stage_persist<1, false>)
                              : (wetdry ? (const void *)k_stage_persist<2, true> : (const void *)k_stage_persist<2, false>);
    const int G = persist_grid(k);
    if (G < 8) return hipErrorInvalidConfiguration;
    MeshDev mm = m;
    PhysDev pp = ph;
    PersistArgs aa = a;
    void *args[] = {(void *)&mm, (void *)&pp, (void *)&aa};
    return hipLaunchCooperativeKernel(k, dim3((unsigned)G), dim3(256), args, 0, s);
}

void launch_node_minmax_end(const MeshDev &m, const PhysDev &ph, const double *means, double *nodemm, int *nodearg,
                            Ctrl *c, int kind, StepRec *rec_end, StepRec *rec_next, bool ties, hipStream_t s)
{
    const int nf = (m.nmem + 7) & ~7;
    const dim3 g((unsigned)(nf + (m.nn + 255) / 256)), b(256);
    if (nodearg)
        hipLaunchKernelGGL(k_node_minmax_end<true>, g, b, 4 * nl_wave_bytes<true>(), s, m, ph, means, nodemm, nodearg, c,
                           kind, nf, rec_end, rec_next, ties ? 1 : 0);
    else
        hipLaunchKernelGGL(k_node_minmax_end<false>, g, b, 4 * nl_wave_bytes<false>(), s, m, ph, means, nodemm, nodearg,
                           c, kind, nf, rec_end, rec_next, ties ? 1 : 0);
}

void launch_step_end(const MeshDev &m, const PhysDev &ph, Ctrl *c, int kind, bool next, hipStream_t s,
                     StepRec *rec_end, StepRec *rec_next, bool ties)
{
    hipLaunchKernelGGL(k_step_end, dim3((unsigned)m.nmem), dim3(256), 0, s, m, ph, c, kind, next ? 1 : 0, rec_end,
                       rec_next, ties ? 1 : 0);
}

void launch_lam_reduce(const MeshDev &m, const PhysDev &ph, Ctrl *c, int slot, int kind, hipStream_t s)
{
    hipLaunchKernelGGL(k_lam_reduce, dim3((unsigned)m.nmem), dim3(256), 0, s, m, ph, c, slot, kind);
}

void launch_step_mark(const MeshDev &m, Ctrl *c, double dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_mark, member_grid(m), dim3(64), 0, s, m, c, dt);
}

void launch_reset_maxlam(const MeshDev &m, Ctrl *c, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset_maxlam, member_grid(m), dim3(64), 0, s, m, c);
}

// ChanNet's channel-junction coupling, junction side (ChanNet/internal_BC.c:
// 16-217), one thread per channel end, its edges in the list's order (the
// order of the reference's sums).  spec: end_off[nends+1], the ends' edges'
// members and member-local edge numbers, end_dir[nends]; vals: end_A, end_Q.
// Writes the edges' bzeta/bqn and ghost[e] (area), ghost[nends+e] (discharge).
__global__ __launch_bounds__(64) void k_couple_ends(MeshDev m, const double *__restrict__ w, double *bzeta,
                                                    double *bqn, int nends, const int *__restrict__ spec,
                                                    const double *__restrict__ vals, double *__restrict__ ghost)
{
    const int e = blockIdx.x * 64 + threadIdx.x;
    if (e >= nends) return;
    const int P = m.pitch;
    const int *end_off = spec;
    const int ne = end_off[nends];
    const int *emem = spec + nends + 1, *eloc = emem + ne, *edir = eloc + ne;
    const int j0 = end_off[e], j1 = end_off[e + 1], n = j1 - j0;
    const double A = vals[e], Q = vals[nends + e];
    const int dir = edir[e];
    double area = 0, qn = 0;
    for (int j = j0; j < j1; ++j) {
        const MemberDev &M = m.mem[emem[j]];
        const int ed = M.ed_off + eloc[j];
        const int4 E = m.edges[ed];
        const int el = E.x, kin = m.einfo[ed] & ED_KIN_MASK;
        const double4 p1 = m.nodes[m.ev[el]], p2 = m.nodes[m.ev[P + el]], p3 = m.nodes[m.ev[2 * P + el]];
        const double xv1 = p1.x, xv2 = p2.x, xv3 = p3.x, yv1 = p1.y, yv2 = p2.y, yv3 = p3.y;
        double x0, x1, y0, y1, len;                                          // internal_BC.c:51-82
        if (kin == 0) {
            x0 = xv1; x1 = xv2; y0 = yv1; y1 = yv2;
            len = sqrt((xv1 - xv2) * (xv1 - xv2) + (yv1 - yv2) * (yv1 - yv2));
        } else if (kin == 1) {
            x0 = xv2; x1 = xv3; y0 = yv2; y1 = yv3;
            len = sqrt((xv2 - xv3) * (xv2 - xv3) + (yv2 - yv3) * (yv2 - yv3));
        } else {
            x0 = xv3; x1 = xv1; y0 = yv3; y1 = yv1;
            len = sqrt((xv1 - xv3) * (xv1 - xv3) + (yv1 - yv3) * (yv1 - yv3));
        }
        const double zeta = w[kin * P + el], qx = w[(3 + kin) * P + el], qy = w[(6 + kin) * P + el];
        const double zj = 0.5 * (m.nodes[E.z].z + m.nodes[E.w].z);            // :87-88
        const double H = zeta + zj;
        const double jac = (xv2 - xv1) * (yv3 - yv1) - (xv3 - xv1) * (yv2 - yv1);
        const int sj = isign(jac);
        const double nx = sj * (y1 - y0) / len;                              // :92-93
        const double ny = sj * (x0 - x1) / len;
        const double qnj = qx * nx + qy * ny;
        area += len * H;                                                     // :97-98
        qn += qnj * len;
        bzeta[ed] = A / (n * len) - zj;                                      // :101-107, :197-203
        bqn[ed] = (dir > 0 ? Q : -Q) / (n * len);
    }
    ghost[e] = area;                                                         // :110-114, :208-214
    ghost[nends + e] = dir > 0 ? qn : -qn;
}

// DG-SWEM's coupling edge with DGSHED's 1-D channels, 2-D side: the
// channel end's values become the edge's boundary data
// (obtain_channels_values, DGSHED/couple_domains.c:248-295, with getH of
// DGSHED/constitutive_equations.c:6-24, as CHANNEL_EDGE_HYDRO reads them,
// src/channel_edge_hydro.F:72-75), and el_in's trace averaged over the
// edge's two Gauss points goes to the channel end's ghost values
// (POPULATE_SWE_ARRAY, src/couple2dgshed.F:22-81; receive_swe_values,
// DGSHED/couple_domains.c:297-335).  The trace uses this element's CR basis
// at the Gauss points as Compute2DInnerProducts evaluates it.  One thread per
// coupled edge, the oracle's orc_couple_dgshed operation for operation.
// spec: members[n], member-local edges[n], conn[n], then flag[n] (written:
// 1 if the edge is no boundary edge, which is left unchanged); vals: chan[6n]
// (A, Q, b, m1, m2, z), geom[4n] (b, m1, m2, z), then swe[2n], ghost[2n]
// (written).
__device__ __forceinline__ double channel_depth(double A, double b, double m1, double m2)
{
    double H;
    if (m1 > 0 || m2 > 0)
        H = (-b + sqrt(b * b + 2 * A * m1 + 2 * A * m2)) / (m1 + m2);
    else
        H = A / b;
    return H;
}

__global__ __launch_bounds__(64) void k_couple_dgshed(MeshDev m, PhysDev ph, const double *__restrict__ w,
                                                      double *bzeta, double *bqn, int n, int *spec,
                                                      double *vals)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const int *emem = spec, *eloc = spec + n, *conn = spec + 2 * n;
    int *flag = spec + 3 * n;
    const double *chan = vals + 6 * i, *gm = vals + 6 * n + 4 * i;
    double *swe = vals + 10 * n, *ghost = vals + 12 * n;
    const MemberDev &M = m.mem[emem[i]];
    const int ed = M.ed_off + eloc[i];
    const int inf = m.einfo[ed];
    if (!(inf & EI_BOUNDARY)) {
        flag[i] = 1;
        return;
    }
    flag[i] = 0;
    const int4 E = m.edges[ed];
    const int el = E.x, kin = inf & ED_KIN_MASK;
    const double4 A = m.nodes[E.z], B = m.nodes[E.w];
    const double len = sqrt((B.x - A.x) * (B.x - A.x) + (B.y - A.y) * (B.y - A.y));   // compute2DL.c:98-132
    const double sg = (double)(((inf >> EI_SGN_SHIFT) & 3) - 1);
    const double nx = sg * (B.y - A.y) / len;
    const double ny = sg * (A.x - B.x) / len;
    // channel -> 2-D
    const double zeta_ch = channel_depth(chan[0], chan[2], chan[3], chan[4]) - chan[5];
    const double qn_ch = conn[i] == 0 ? chan[1] : -chan[1];
    // 2-D -> channel: the trace at the Gauss points of edge kin
    GeoRegs G;
    DivExact D;
    geometry_of(m, el, ph.g, [&](int j, double v) { G.v[j] = v; }, D);
    const int P = m.pitch;
    double ze_avg = 0.0, q_n_sum = 0.0;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const double P1 = G(GX_PS + 2 * kin + q), P2 = G(GX_PS + 6 + 2 * kin + q), P3 = G(GX_PS + 12 + 2 * kin + q);
        double tr[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            double t = w[(3 * f) * P + el] * P1;
            t = t + w[(3 * f + 1) * P + el] * P2;
            t = t + w[(3 * f + 2) * P + el] * P3;
            tr[f] = t;
        }
        const double q_n_int = tr[1] * nx + tr[2] * ny;
        ze_avg = ze_avg + tr[0] * 1.0;
        q_n_sum = q_n_sum + q_n_int * 1.0;
    }
    const double sz = 0.5 * ze_avg, sq = 0.5 * len * q_n_sum;
    swe[2 * i] = sz;
    swe[2 * i + 1] = sq;
    const double htval = sz + gm[3];
    ghost[2 * i] = htval * gm[0] + 0.5 * (gm[1] + gm[2]) * htval * htval;
    ghost[2 * i + 1] = conn[i] == 0 ? sq : -sq;
    bzeta[ed] = zeta_ch;
    bqn[ed] = qn_ch / len;
}

void launch_couple_dgshed(const MeshDev &m, const PhysDev &ph, const double *w, double *bzeta, double *bqn, int n,
                          int *spec, double *vals, hipStream_t s)
{
    hipLaunchKernelGGL(k_couple_dgshed, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, m, ph, w, bzeta, bqn, n,
                       spec, vals);
}

void launch_couple_ends(const MeshDev &m, const double *w, double *bzeta, double *bqn, int nends, const int *spec,
                        const double *vals, double *ghost, hipStream_t s)
{
    hipLaunchKernelGGL(k_couple_ends, dim3((unsigned)((nends + 63) / 64)), dim3(64), 0, s, m, w, bzeta, bqn, nends,
                       spec, vals, ghost);
}

void launch_math_selftest(int n, const double *x, const double *y, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_math_selftest, grid_for(n), dim3(256), 0, s, n, x, y, out);
}

// the reverse sweep (shares the device functions above)
#include "dgr_adjoint_kernels.h"

}  // namespace dgr
