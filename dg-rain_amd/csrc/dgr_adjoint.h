// dgr_adjoint.h -- hand-written reverse-mode derivatives of the discrete 2D
// p=1 DG step's pieces, for the gradient of a loss with respect to the state,
// nodal Manning n and nodal depth z.
//
// Each *_adj function re-evaluates the forward operation from its inputs and
// accumulates  xbar += (d out / d x)^T outbar  for every input x.  The forward
// operations are the ones of dgr_kernels.hip (numericalFlux2D.c:6-108,
// 2DInnerProducts.c:8-570, compute2DL.c:338-411, SlopeLimiter2D.c:157-343);
// their non-smooth parts (max, |.|, min/max clipping, limiter branches) are
// differentiated on the branch the forward evaluation takes, which is the
// forward pass's active set.  Derivatives need not reproduce the forward's
// rounding, so algebraically equal forms are used where they are simpler.
//
// Host and device: the same functions are compiled into the GPU kernels and
// into the host-side finite-difference checks.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "dgr_device.h"
#include "dgr_math.h"

namespace dgr {

#define DGR_HD __host__ __device__ __forceinline__

DGR_HD double sgn_d(double a) { return a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0); }

// ---------------------------------------------------------------------------
// Roe flux (numericalFlux2D.c:6-108): F[3] and max|lambda| from
// (zi, ze, qxi, qxe, qyi, qye, zb) given (nx, ny, lg, g).  In[7] holds the
// inputs in that order; inbar[7] receives the accumulated adjoints.
// ---------------------------------------------------------------------------
struct RoeIn {
    double zi, ze, qxi, qxe, qyi, qye, zb;
};

DGR_HD double roe_eval(double F[3], const RoeIn &a, double nx, double ny, double lg, double g)
{
    const double Hi = a.zi + a.zb, He = a.ze + a.zb;
    const double ui = a.qxi / Hi, ue = a.qxe / He, vi = a.qyi / Hi, ve = a.qye / He;
    const double j0 = a.zi - a.ze, j1 = a.qxi - a.qxe, j2 = a.qyi - a.qye;
    const double fni0 = (Hi * ui) * nx + (Hi * vi) * ny, fne0 = (He * ue) * nx + (He * ve) * ny;
    const double fni1 = (Hi * ui * ui + 0.5 * lg * (Hi * Hi - a.zb * a.zb)) * nx + (Hi * ui * vi) * ny;
    const double fne1 = (He * ue * ue + 0.5 * lg * (He * He - a.zb * a.zb)) * nx + (He * ue * ve) * ny;
    const double fni2 = (Hi * ui * vi) * nx + (Hi * vi * vi + 0.5 * lg * (Hi * Hi - a.zb * a.zb)) * ny;
    const double fne2 = (He * ue * ve) * nx + (He * ve * ve + 0.5 * lg * (He * He - a.zb * a.zb)) * ny;
    const double Hh = 0.5 * (Hi + He);
    const double sHi = sqrt(Hi), sHe = sqrt(He), S = sHi + sHe;
    const double uh = (ui * sHi + ue * sHe) / S, vh = (vi * sHi + ve * sHe) / S;
    const double c = sqrt(g * Hh);
    const double den = -2 * c;
    const double Q[3][3] = {{(-c - nx * uh - ny * vh) / den, nx / den, ny / den},
                            {(-2 * c * ny * uh + 2 * c * nx * vh) / den, 2 * c * ny / den, -2 * c * nx / den},
                            {(-c + nx * uh + ny * vh) / den, -nx / den, -ny / den}};
    const double sgi = sqrt(g * Hi), sge = sqrt(g * He);
    const double li[3] = {ui * nx + vi * ny - sgi, ui * nx + vi * ny, ui * nx + vi * ny + sgi};
    const double le[3] = {ue * nx + ve * ny - sge, ue * nx + ve * ny, ue * nx + ve * ny + sge};
    const double lh[3] = {uh * nx + vh * ny - c, uh * nx + vh * ny, uh * nx + vh * ny + c};
    double lam[3];
    for (int k = 0; k < 3; ++k) {
        double e = fmax(0.0, lh[k] - li[k]);
        e = fmax(e, le[k] - lh[k]);
        lam[k] = fmax(e, fabs(lh[k]));
    }
    double p1[3], p2[3];
    for (int r = 0; r < 3; ++r) p1[r] = Q[r][0] * j0 + Q[r][1] * j1 + Q[r][2] * j2;
    for (int r = 0; r < 3; ++r) p2[r] = lam[r] * p1[r];
    const double R1[3] = {uh - c * nx, -ny, uh + c * nx}, R2[3] = {vh - c * ny, nx, vh + c * ny};
    F[0] = 0.5 * (fni0 + fne0) + 0.5 * (p2[0] + p2[2]);
    F[1] = 0.5 * (fni1 + fne1) + 0.5 * (R1[0] * p2[0] + R1[1] * p2[1] + R1[2] * p2[2]);
    F[2] = 0.5 * (fni2 + fne2) + 0.5 * (R2[0] * p2[0] + R2[1] * p2[1] + R2[2] * p2[2]);
    double mx = fabs(li[0]);
    const double ops[5] = {fabs(le[0]), fabs(li[1]), fabs(le[1]), fabs(li[2]), fabs(le[2])};
    for (int i = 0; i < 5; ++i) mx = ops[i] > mx ? ops[i] : mx;
    return mx;
}

// Reverse of roe_eval: Fb[3] and lamb (adjoint of the returned max|lambda|)
// accumulate into ib.  With tie, lamb applies only if the max|lambda| that
// roe_eval would return (the same operands, the same fold) equals target.
DGR_HD void roe_adj_impl(const RoeIn &a, double nx, double ny, double lg, double g, const double Fb[3], double lamb,
                         bool tie, double target, RoeIn &ib)
{
    // forward values
    const double Hi = a.zi + a.zb, He = a.ze + a.zb;
    const double ui = a.qxi / Hi, ue = a.qxe / He, vi = a.qyi / Hi, ve = a.qye / He;
    const double j[3] = {a.zi - a.ze, a.qxi - a.qxe, a.qyi - a.qye};
    const double Hh = 0.5 * (Hi + He);
    const double sHi = sqrt(Hi), sHe = sqrt(He), S = sHi + sHe;
    const double uh = (ui * sHi + ue * sHe) / S, vh = (vi * sHi + ve * sHe) / S;
    const double c = sqrt(g * Hh);
    const double den = -2 * c;
    const double num[3] = {-c - nx * uh - ny * vh, -2 * c * ny * uh + 2 * c * nx * vh, -c + nx * uh + ny * vh};
    // Q only weighs derivatives: reciprocal products
    const double rden = recip_d(den);
    const double Q[3][3] = {{num[0] * rden, nx * rden, ny * rden},
                            {num[1] * rden, 2 * c * ny * rden, -2 * c * nx * rden},
                            {num[2] * rden, -nx * rden, -ny * rden}};
    const double sgi = sqrt(g * Hi), sge = sqrt(g * He);
    const double li[3] = {ui * nx + vi * ny - sgi, ui * nx + vi * ny, ui * nx + vi * ny + sgi};
    const double le[3] = {ue * nx + ve * ny - sge, ue * nx + ve * ny, ue * nx + ve * ny + sge};
    const double lh[3] = {uh * nx + vh * ny - c, uh * nx + vh * ny, uh * nx + vh * ny + c};
    if (tie) {
        // roe_eval's max|lambda| fold over the same operands
        double mx = fabs(li[0]);
        const double ops[5] = {fabs(le[0]), fabs(li[1]), fabs(le[1]), fabs(li[2]), fabs(le[2])};
        for (int i = 0; i < 5; ++i) mx = ops[i] > mx ? ops[i] : mx;
        if (!(mx == target)) lamb = 0.0;
    }
    double lam[3];
    int act[3];   // 0: zero, 1: lh-li, 2: le-lh, 3: |lh|
    for (int k = 0; k < 3; ++k) {
        double e = 0.0;
        int ae = 0;
        const double t1 = lh[k] - li[k];
        if (t1 > e) { e = t1; ae = 1; }
        const double t2 = le[k] - lh[k];
        if (t2 > e) { e = t2; ae = 2; }
        const double t3 = fabs(lh[k]);
        if (t3 > e) { e = t3; ae = 3; }
        lam[k] = e;
        act[k] = ae;
    }
    double p1[3], p2[3];
    for (int r = 0; r < 3; ++r) p1[r] = Q[r][0] * j[0] + Q[r][1] * j[1] + Q[r][2] * j[2];
    for (int r = 0; r < 3; ++r) p2[r] = lam[r] * p1[r];
    const double R1[3] = {uh - c * nx, -ny, uh + c * nx}, R2[3] = {vh - c * ny, nx, vh + c * ny};

    // reverse
    double Hib = 0, Heb = 0, uib = 0, ueb = 0, vib = 0, veb = 0, zbb = 0;
    double uhb = 0, vhb = 0, cb = 0, sgib = 0, sgeb = 0;
    // F = 0.5 (fni + fne) + 0.5 p3
    const double fnb[3] = {0.5 * Fb[0], 0.5 * Fb[1], 0.5 * Fb[2]};
    const double p3b[3] = {0.5 * Fb[0], 0.5 * Fb[1], 0.5 * Fb[2]};
    // p3 = [1 0 1; R1; R2] p2
    double p2b[3];
    for (int k = 0; k < 3; ++k) p2b[k] = R1[k] * p3b[1] + R2[k] * p3b[2];
    p2b[0] += p3b[0];
    p2b[2] += p3b[0];
    const double R1b0 = p2[0] * p3b[1], R1b2 = p2[2] * p3b[1];
    const double R2b0 = p2[0] * p3b[2], R2b2 = p2[2] * p3b[2];
    uhb += R1b0 + R1b2;
    cb += -nx * R1b0 + nx * R1b2;
    vhb += R2b0 + R2b2;
    cb += -ny * R2b0 + ny * R2b2;
    // p2 = lam * p1
    double p1b[3], lamb_k[3];
    for (int k = 0; k < 3; ++k) { p1b[k] = lam[k] * p2b[k]; lamb_k[k] = p1[k] * p2b[k]; }
    // p1 = Q j
    double jb[3] = {0, 0, 0}, Qb[3][3];
    for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) { Qb[r][s] = p1b[r] * j[s]; jb[s] += Q[r][s] * p1b[r]; }
    // entropy fix
    double lib[3] = {0, 0, 0}, leb[3] = {0, 0, 0}, lhb[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        if (act[k] == 1) { lhb[k] += lamb_k[k]; lib[k] -= lamb_k[k]; }
        else if (act[k] == 2) { leb[k] += lamb_k[k]; lhb[k] -= lamb_k[k]; }
        else if (act[k] == 3) lhb[k] += sgn_d(lh[k]) * lamb_k[k];
    }
    // max |lambda| (left fold, ties keep the earlier operand)
    // (an exact tie shares the adjoint among the tied operands, the value a
    // central difference gives for |x| at a kink such as l1 = -l3 at rest)
    if (lamb != 0.0) {
        const double vals[6] = {li[0], le[0], li[1], le[1], li[2], le[2]};
        double mx = 0.0;
        for (int i = 0; i < 6; ++i) mx = fabs(vals[i]) > mx ? fabs(vals[i]) : mx;
        int nt = 0;
        for (int i = 0; i < 6; ++i) nt += fabs(vals[i]) == mx;
        for (int i = 0; i < 6; ++i)
            if (fabs(vals[i]) == mx) {
                const double d = sgn_d(vals[i]) * lamb / nt;
                if (i & 1) leb[i >> 1] += d; else lib[i >> 1] += d;
            }
    }
    // eigenvalues
    uib += nx * (lib[0] + lib[1] + lib[2]);
    vib += ny * (lib[0] + lib[1] + lib[2]);
    sgib += -lib[0] + lib[2];
    ueb += nx * (leb[0] + leb[1] + leb[2]);
    veb += ny * (leb[0] + leb[1] + leb[2]);
    sgeb += -leb[0] + leb[2];
    uhb += nx * (lhb[0] + lhb[1] + lhb[2]);
    vhb += ny * (lhb[0] + lhb[1] + lhb[2]);
    cb += -lhb[0] + lhb[2];
    // the reverse sweep multiplies by reciprocals (the forward values above
    // keep the reference's divisions: they decide the active branches)
    Hib += sgib * ((0.5 * g) * recip_d(sgi));
    Heb += sgeb * ((0.5 * g) * recip_d(sge));
    // Q = num / den and the constant-numerator entries
    double denb = 0;
    for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) denb += -Q[r][s] * Qb[r][s];
    denb *= rden;
    const double n0b = Qb[0][0] * rden, n1b = Qb[1][0] * rden, n2b = Qb[2][0] * rden;
    cb += -n0b;
    uhb += -nx * n0b;
    vhb += -ny * n0b;
    cb += n1b * (-2 * ny * uh + 2 * nx * vh);
    uhb += n1b * (-2 * c * ny);
    vhb += n1b * (2 * c * nx);
    cb += -n2b;
    uhb += nx * n2b;
    vhb += ny * n2b;
    cb += (2 * ny * Qb[1][1] - 2 * nx * Qb[1][2]) * rden;
    cb += -2 * denb;
    // c = sqrt(g Hh)
    const double Hhb = cb * ((0.5 * g) * recip_d(c));
    Hib += 0.5 * Hhb;
    Heb += 0.5 * Hhb;
    // Roe averages
    const double rS = recip_d(S);
    const double numub = uhb * rS, numvb = vhb * rS;
    const double Sb = -(uh * uhb + vh * vhb) * rS;
    uib += sHi * numub;
    ueb += sHe * numub;
    vib += sHi * numvb;
    veb += sHe * numvb;
    const double sHib = ui * numub + vi * numvb + Sb, sHeb = ue * numub + ve * numvb + Sb;
    Hib += sHib * (0.5 * recip_d(sHi));
    Heb += sHeb * (0.5 * recip_d(sHe));
    // physical normal fluxes
    {
        const double f1x = nx * fnb[0], f1y = ny * fnb[0], f2x = nx * fnb[1], f2y = ny * fnb[1];
        const double f3x = nx * fnb[2], f3y = ny * fnb[2];
        // in side
        Hib += ui * f1x + vi * f1y + (ui * ui + lg * Hi) * f2x + ui * vi * f2y + ui * vi * f3x + (vi * vi + lg * Hi) * f3y;
        uib += Hi * f1x + 2 * Hi * ui * f2x + Hi * vi * f2y + Hi * vi * f3x;
        vib += Hi * f1y + Hi * ui * f2y + Hi * ui * f3x + 2 * Hi * vi * f3y;
        zbb += -lg * a.zb * (f2x + f3y);
        // ex side
        Heb += ue * f1x + ve * f1y + (ue * ue + lg * He) * f2x + ue * ve * f2y + ue * ve * f3x + (ve * ve + lg * He) * f3y;
        ueb += He * f1x + 2 * He * ue * f2x + He * ve * f2y + He * ve * f3x;
        veb += He * f1y + He * ue * f2y + He * ue * f3x + 2 * He * ve * f3y;
        zbb += -lg * a.zb * (f2x + f3y);
    }
    // velocities
    const double rHi = recip_d(Hi), rHe = recip_d(He);
    ib.qxi += uib * rHi;
    ib.qyi += vib * rHi;
    Hib += -(ui * uib + vi * vib) * rHi;
    ib.qxe += ueb * rHe;
    ib.qye += veb * rHe;
    Heb += -(ue * ueb + ve * veb) * rHe;
    // jumps
    ib.zi += jb[0];
    ib.ze -= jb[0];
    ib.qxi += jb[1];
    ib.qxe -= jb[1];
    ib.qyi += jb[2];
    ib.qye -= jb[2];
    // depths
    ib.zi += Hib;
    ib.ze += Heb;
    ib.zb += Hib + Heb + zbb;
}

DGR_HD void roe_adj(const RoeIn &a, double nx, double ny, double lg, double g, const double Fb[3], double lamb,
                    RoeIn &ib)
{
    roe_adj_impl(a, nx, ny, lg, g, Fb, lamb, false, 0.0, ib);
}

// the tie-gated form: lamb is shared only if this evaluation's max|lambda|
// equals target (the recorded max of the stage)
DGR_HD void roe_adj_tie(const RoeIn &a, double nx, double ny, double lg, double g, const double Fb[3], double lamb,
                        double target, RoeIn &ib)
{
    roe_adj_impl(a, nx, ny, lg, g, Fb, lamb, true, target, ib);
}

// ---------------------------------------------------------------------------
// RT0 fit (2DInnerProducts.c:134-146) as the linear map (R1,R2,R3) -> (a,b,c)
// and its transpose.
// ---------------------------------------------------------------------------
struct Rt0 {
    double A[3], B[3], C[3];   // a = A.R, b = B.R, c = C.R
};

DGR_HD Rt0 rt0_map(double M11, double M12, double M13, double M21, double M22, double M23, double M31, double M32,
                   double M33, double den)
{
    Rt0 m;
    const double r = recip_d(den);
    m.A[0] = (M23 * M32 - M22 * M33) * r;
    m.A[1] = (M12 * M33 - M13 * M32) * r;
    m.A[2] = (M13 * M22 - M12 * M23) * r;
    m.B[0] = -(M23 * M31 - M21 * M33) * r;
    m.B[1] = -(M11 * M33 - M13 * M31) * r;
    m.B[2] = -(M13 * M21 - M11 * M23) * r;
    m.C[0] = (M22 * M31 - M21 * M32) * r;
    m.C[1] = (M11 * M32 - M12 * M31) * r;
    m.C[2] = (M12 * M21 - M11 * M22) * r;
    return m;
}

// ---------------------------------------------------------------------------
// Element RHS (2DInnerProducts.c:14-566 + compute2DL.c:344-409) reverse.
// Geometry quantities come from geometry_of (GeoIdx); the ones that depend on
// the nodal depths are the midpoint depths mz (GX_MZ) and the bed-slope terms
// (GX_BGX/BGY), the one that depends on Manning n is GX_GNN.
// Inputs: W[9] state, wd_self; outputs adjoint Rb[9].
// Accumulates Wb[9], Fhb[9], mzb[3], gnnb[3], bgxb, bgyb.
// ---------------------------------------------------------------------------
// The reverse in two parts that read different geometry: the surface
// integrals (-> Fhb) and the volume, friction and bed terms (-> Wb, mzb,
// gnnb, bgxb, bgyb).  A kernel may evaluate the geometry separately for
// each, so that only one part's geometry is live at a time.
template <typename Geo>
DGR_HD Rt0 rt0_map_of(const Geo &G)
{
    return rt0_map(G(GX_NX), G(GX_M2), G(GX_NY), G(GX_NX + 1), G(GX_M2 + 1), G(GX_NY + 1), G(GX_NX + 2),
                   G(GX_M2 + 2), G(GX_NY + 2), G(GX_DEN));
}

template <typename Geo>
DGR_HD void element_rhs_adj_surface(const Geo &G, const double Rb[9], double Fhb[9])
{
    const double nxs[3] = {G(GX_NX), G(GX_NX + 1), G(GX_NX + 2)};
    const double nys[3] = {G(GX_NY), G(GX_NY + 1), G(GX_NY + 2)};
    const Rt0 T = rt0_map_of(G);
    const double Mi = G(GX_MI);
    // R = Mi (..) : adjoints of ze, qe, re
    double eb[9];
    for (int i = 0; i < 9; ++i) eb[i] = Mi * Rb[i];
    // surface term SI = -..: SIb = -eb
    // SI[f*3+b] = sum_ed ej (Fn[f][ed][0] p0 + Fn[f][ed][1] p1),
    // Fn = (a + b xq) nx + (c + b yq) ny
    double ab[3] = {0, 0, 0}, bb_[3] = {0, 0, 0}, cb_[3] = {0, 0, 0};
    for (int ed = 0; ed < 3; ++ed) {
        const double ej = G(GX_EJ + ed);
        for (int n = 0; n < 2; ++n) {
            const double xq = G(GX_XE + 2 * ed + n), yq = G(GX_YE + 2 * ed + n);
            for (int f = 0; f < 3; ++f) {
                double Fnb = 0;
                for (int b = 0; b < 3; ++b) Fnb += ej * G(GX_PS + 6 * b + 2 * ed + n) * (-eb[f * 3 + b]);
                ab[f] += nxs[ed] * Fnb;
                cb_[f] += nys[ed] * Fnb;
                bb_[f] += (xq * nxs[ed] + yq * nys[ed]) * Fnb;
            }
        }
    }
    for (int f = 0; f < 3; ++f)
        for (int k = 0; k < 3; ++k) Fhb[f * 3 + k] += T.A[k] * ab[f] + T.B[k] * bb_[f] + T.C[k] * cb_[f];
}

template <typename Geo>
DGR_HD void element_rhs_adj_volume(const Geo &G, const double W[9], int wd_self, double g, const double Rb[9],
                                   double Wb[9], double mzb[3], double gnnb[3], double &bgxb, double &bgyb)
{
    const double nxs[3] = {G(GX_NX), G(GX_NX + 1), G(GX_NX + 2)};
    const double nys[3] = {G(GX_NY), G(GX_NY + 1), G(GX_NY + 2)};
    const Rt0 T = rt0_map_of(G);
    const double Mi = G(GX_MI), aw = G(GX_AW), aj = G(GX_AJ), bgx = G(GX_BGX), bgy = G(GX_BGY);
    const double weight = 1.0 / 6;
    const double lg = (wd_self == 0) ? 0.0 : g;
    double eb[9];
    for (int i = 0; i < 9; ++i) eb[i] = Mi * Rb[i];
    // volume term VF[f][b] = aj (0.5 va GPx_b + 0.5 vc GPy_b + vb weight GS_b), (va,vb,vc) = rt0(Fnp)
    double Fnpb[3][3];   // [field][k]
    for (int f = 0; f < 3; ++f) {
        double vab = 0, vbb = 0, vcb = 0;
        for (int b = 0; b < 3; ++b) {
            const double x = eb[f * 3 + b];
            vab += aj * 0.5 * G(GX_GP + 2 * b) * x;
            vcb += aj * 0.5 * G(GX_GP + 2 * b + 1) * x;
            vbb += aj * weight * G(GX_GS + b) * x;
        }
        for (int k = 0; k < 3; ++k) Fnpb[f][k] = T.A[k] * vab + T.B[k] * vbb + T.C[k] * vcb;
    }
    // point fluxes, friction and bed terms per midpoint k
    for (int k = 0; k < 3; ++k) {
        const double mz = G(GX_MZ + k), gnn = G(GX_GNN + k);
        const double zr = W[k], qx = W[3 + k], qy = W[6 + k];
        const double H = zr + mz;
        const double rH = recip_d(H);
        const double u = qx * rH, v = qy * rH;
        const double sp = sqrt(u * u + v * v);
        const double rH43 = recip_d(H * cbrt(H));
        const double tau = gnn * sp * rH43;
        double Hb = 0, qxb = 0, qyb = 0, zrb = 0, mzb_explicit = 0;
        // Fx/Fy adjoints from the normal point fluxes
        const double fxb[3] = {nxs[k] * Fnpb[0][k], nxs[k] * Fnpb[1][k], nxs[k] * Fnpb[2][k]};
        const double fyb[3] = {nys[k] * Fnpb[0][k], nys[k] * Fnpb[1][k], nys[k] * Fnpb[2][k]};
        // Fx0 = qx, Fy0 = qy
        qxb += fxb[0];
        qyb += fyb[0];
        // Fx1 = qx^2/H + 0.5 lg (H^2 - mz^2)
        qxb += 2 * u * fxb[1];
        Hb += (-u * u + lg * H) * fxb[1];
        mzb_explicit += -lg * mz * fxb[1];
        // Fx2 = Fy1 = qx qy / H
        const double x2 = fxb[2] + fyb[1];
        qxb += v * x2;
        qyb += u * x2;
        Hb += -u * v * x2;
        // Fy2 = qy^2/H + 0.5 lg (H^2 - mz^2)
        qyb += 2 * v * fyb[2];
        Hb += (-v * v + lg * H) * fyb[2];
        mzb_explicit += -lg * mz * fyb[2];
        // friction fx = aw tau qx, fy = aw tau qy enter with a minus sign
        const double qeb = eb[3 + k], reb = eb[6 + k];
        const double taub = -aw * (qx * qeb + qy * reb);
        qxb += -aw * tau * qeb;
        qyb += -aw * tau * reb;
        // tau = gnn sp / H^(4/3)
        gnnb[k] += sp * rH43 * taub;
        const double spb = gnn * rH43 * taub;
        Hb += -(4.0 / 3.0) * tau * rH * taub;
        if (sp > 0) {
            const double rsp = spb / sp;
            const double ub = u * rsp, vb = v * rsp;
            qxb += ub * rH;
            qyb += vb * rH;
            Hb += -(u * ub + v * vb) * rH;
        }
        // bed slope terms bx = bgx zeta, by = bgy zeta (zero when dry)
        if (wd_self != 0) {
            zrb += bgx * qeb + bgy * reb;
            bgxb += zr * qeb;
            bgyb += zr * reb;
        }
        // H = zr + mz
        zrb += Hb;
        mzb[k] += Hb + mzb_explicit;
        Wb[k] += zrb;
        Wb[3 + k] += qxb;
        Wb[6 + k] += qyb;
    }
}

template <typename Geo>
DGR_HD void element_rhs_adj(const Geo &G, const double W[9], int wd_self, double g, const double Rb[9], double Wb[9],
                            double Fhb[9], double mzb[3], double gnnb[3], double &bgxb, double &bgyb)
{
    element_rhs_adj_surface(G, Rb, Fhb);
    element_rhs_adj_volume(G, W, wd_self, g, Rb, Wb, mzb, gnnb, bgxb, bgyb);
}

// forward decisions of one field's limiter pass (SlopeLimiter2D.c:166-303):
// the clamp branch per vertex, the sign of each redistribution pass, which
// vertices were positive in each pass and which took the uncapped
// reduction (bit 3*pass + k), and the final vertex values
struct LimFieldDec {
    int clampsel[3];   // 0: V, 1: Vmax, 2: Vmin
    int sd[3];         // per pass
    unsigned pos, use0;
    double Vfin[3];
};

DGR_HD void limit_field_forward(const double S[9], int iv, const double Vmin[3], const double Vmax[3],
                                LimFieldDec &d)
{
    const double w1 = S[3 * iv], w2 = S[3 * iv + 1], w3 = S[3 * iv + 2];
    double V[3] = {w1 - w2 + w3, w1 + w2 - w3, -w1 + w2 + w3};
    const double avg = (w1 + w2 + w3) / 3;
    for (int k = 0; k < 3; ++k) {
        // at an exact tie the element's own value is taken as the active
        // branch: the clip is then inactive, as it is in both directions
        // of a central difference when several elements share the bound
        // (lake at rest), and no adjoint is routed through a tied bound
        const double lo = (V[k] < Vmax[k]) ? V[k] : Vmax[k];
        int sel = (V[k] <= Vmax[k]) ? 0 : 1;
        if (Vmin[k] > lo) sel = 2;
        d.clampsel[k] = sel;
        V[k] = fmax(lo, Vmin[k]);
    }
    d.pos = 0;
    d.use0 = 0;
    for (int ll = 0; ll < 3; ++ll) {
        const double sumloc = (V[0] + V[1] + V[2]) / 3;
        double sumdiff = (sumloc - avg) * 3;
        const int s = sumdiff < 0 ? -1 : (sumdiff == 0 ? 0 : 1);
        d.sd[ll] = s;
        const double diff[3] = {(V[0] - avg) * s, (V[1] - avg) * s, (V[2] - avg) * s};
        double kdp = (double)((diff[0] > 0) + (diff[1] > 0) + (diff[2] > 0));
        for (int k = 0; k < 3; ++k) {
            double redfac;
            if (diff[k] > 0) {
                d.pos |= 1u << (3 * ll + k);
                // sumdiff*s / max(1, kdp), kdp >= 1 here: x/1 = x and x/2 =
                // x*0.5 exactly, and only k = 0 can see kdp = 3 (all three
                // positive)
                const double num = sumdiff * s;
                redfac = (k == 0 && kdp > 2.0) ? num / 3 : (kdp > 1.0) ? num * 0.5 : num;
                kdp -= 1;
            } else {
                redfac = 0;
            }
            const double redmax = (s > 0) ? V[k] - Vmin[k] : Vmax[k] - V[k];
            if (redfac < redmax) d.use0 |= 1u << (3 * ll + k);
            redfac = (redfac < redmax) ? redfac : redmax;
            sumdiff -= redfac * s;
            V[k] -= redfac * s;
        }
    }
    for (int k = 0; k < 3; ++k) d.Vfin[k] = V[k];
}

// a field pass's decisions (all the reverse reads) in one word: clamp
// selections (bits 0-5), pass signs + 1 (6-11), positive diffs (12-20),
// redistributions that kept redfac (21-29)
DGR_HD unsigned lim_dec_pack(const LimFieldDec &d)
{
    unsigned p = 0;
    for (int k = 0; k < 3; ++k) p |= (unsigned)d.clampsel[k] << (2 * k);
    for (int ll = 0; ll < 3; ++ll) p |= (unsigned)(d.sd[ll] + 1) << (6 + 2 * ll);
    return p | (d.pos << 12) | (d.use0 << 21);
}

DGR_HD LimFieldDec lim_dec_unpack(unsigned p)
{
    LimFieldDec d;
    for (int k = 0; k < 3; ++k) d.clampsel[k] = (int)((p >> (2 * k)) & 3u);
    for (int ll = 0; ll < 3; ++ll) d.sd[ll] = (int)((p >> (6 + 2 * ll)) & 3u) - 1;
    d.pos = (p >> 12) & 0x1ffu;
    d.use0 = (p >> 21) & 0x1ffu;
    for (int k = 0; k < 3; ++k) d.Vfin[k] = 0.0;
    return d;
}

// reverse of one field's pass given its decisions: Sb the adjoint of the
// limited midpoints; accumulates Sin_b and the bound adjoints
DGR_HD void limit_field_reverse(const LimFieldDec &d, int iv, const double Sb[9], double Sin_b[9], double Vminb[3],
                                double Vmaxb[3])
{
    // mids: S_out = ((V0+V1)/2, (V1+V2)/2, (V2+V0)/2)
    double Vb[3];
    Vb[0] = 0.5 * (Sb[3 * iv] + Sb[3 * iv + 2]);
    Vb[1] = 0.5 * (Sb[3 * iv] + Sb[3 * iv + 1]);
    Vb[2] = 0.5 * (Sb[3 * iv + 1] + Sb[3 * iv + 2]);
    double avgb = 0;
    for (int ll = 2; ll >= 0; --ll) {
        const int s = d.sd[ll];
        double sumdiffb = 0;   // adjoint of the running sumdiff after the k loop
        for (int k = 2; k >= 0; --k) {
            // V_k -= red*s ; sumdiff -= red*s
            const double redb = -s * Vb[k] - s * sumdiffb;
            const unsigned bit = 1u << (3 * ll + k);
            if (d.use0 & bit) {
                if (d.pos & bit) {
                    // the divisor max(1, kdp): positives at k and after in this pass
                    const unsigned later = (d.pos >> (3 * ll + k)) & (7u >> k);
                    const int cnt = (int)(later & 1u) + (int)((later >> 1) & 1u) + (int)((later >> 2) & 1u);
                    const double x = redb * s;
                    sumdiffb += (cnt == 3) ? x / 3 : (cnt == 2) ? x * 0.5 : x;
                }
            } else {
                if (s > 0) { Vb[k] += redb; Vminb[k] -= redb; }
                else { Vmaxb[k] += redb; Vb[k] -= redb; }
            }
        }
        // sumdiff = (sum V / 3 - avg) * 3 = sum V - 3 avg
        for (int k = 0; k < 3; ++k) Vb[k] += sumdiffb;
        avgb += -3 * sumdiffb;
    }
    // clamp
    double V0b[3];
    for (int k = 0; k < 3; ++k) {
        V0b[k] = 0;
        if (d.clampsel[k] == 0) V0b[k] = Vb[k];
        else if (d.clampsel[k] == 1) Vmaxb[k] += Vb[k];
        else Vminb[k] += Vb[k];
    }
    // V0 = (w1-w2+w3, w1+w2-w3, -w1+w2+w3), avg = (w1+w2+w3)/3
    const double avgb3 = avgb / 3;
    Sin_b[3 * iv] += V0b[0] + V0b[1] - V0b[2] + avgb3;
    Sin_b[3 * iv + 1] += -V0b[0] + V0b[1] + V0b[2] + avgb3;
    Sin_b[3 * iv + 2] += V0b[0] - V0b[1] + V0b[2] + avgb3;
}

// ---------------------------------------------------------------------------
// Limiter element pass (SlopeLimiter2D.c:157-343) reverse for one element:
// S[9] is the element's state before limiting, Vmin/Vmax[f][k] the node
// bounds at its vertices, zv[3] the vertex depths; Sb[9] the adjoint of the
// limited state.  Returns the modify flag.  Accumulates Sin_b[9] (into the
// element's own pre-limit state) and Vminb/Vmaxb[f][k]; Sout (optional)
// receives the limited state.  The zeta pass decides modify; each field is then
// replayed and reversed on its own, so one field's decisions are live at a
// time.
// ---------------------------------------------------------------------------
DGR_HD int limit_adj(const double S[9], const double Vmin[3][3], const double Vmax[3][3], const double zv[3],
                     double H0, const double Sb[9], double Sin_b[9], double Vminb[3][3], double Vmaxb[3][3],
                     double Sout[9] = nullptr)
{
    LimFieldDec d;
    limit_field_forward(S, 0, Vmin[0], Vmax[0], d);
    int modify = 0;
    {
        const double w1 = S[0], w2 = S[1], w3 = S[2];
        const double V0 = w1 - w2 + w3, V1 = w1 + w2 - w3, V2 = -w1 + w2 + w3;
        const double Ha = V0 + zv[0], Hb_ = V1 + zv[1], Hc = V2 + zv[2];
        const double Hm0 = Vmin[0][0] + zv[0], Hm1 = Vmin[0][1] + zv[1], Hm2 = Vmin[0][2] + zv[2];
        if (Ha > H0 && Hb_ > H0 && Hc > H0 && Hm0 > H0 && Hm1 > H0 && Hm2 > H0) modify = 1;
        const double Fa = d.Vfin[0] + zv[0], Fb = d.Vfin[1] + zv[1], Fc = d.Vfin[2] + zv[2];
        if (Fa < H0 || Fb < H0 || Fc < H0) modify = 0;
    }
    if (!modify) {
        if (Sout)
            for (int i = 0; i < 9; ++i) Sout[i] = S[i];
        for (int i = 0; i < 9; ++i) Sin_b[i] += Sb[i];
        return 0;
    }
    for (int iv = 0; iv < 3; ++iv) {
        if (iv > 0) limit_field_forward(S, iv, Vmin[iv], Vmax[iv], d);
        if (Sout) {
            Sout[3 * iv] = (d.Vfin[0] + d.Vfin[1]) / 2;
            Sout[3 * iv + 1] = (d.Vfin[1] + d.Vfin[2]) / 2;
            Sout[3 * iv + 2] = (d.Vfin[2] + d.Vfin[0]) / 2;
        }
        limit_field_reverse(d, iv, Sb, Sin_b, Vminb[iv], Vmaxb[iv]);
    }
    return 1;
}

// ---------------------------------------------------------------------------
// PDop2D (wetDry2D.c:66-256) reverse for one element, on the branch the
// forward takes.  S[9] is the element's state before PDop (the limited
// state), zv[3] its vertex depths, Sb[9] the adjoint of PDop's output.
// Writes Sin_b[9], the adjoint of S, and accumulates zvb[3] (the depth
// enters through the midpoint bed bz_k = (zv_k + zv_k+1)/2).  Returns 1 if
// PDop changed the element, else 0 (Sin_b = Sb).
//
// The forward, with T the midpoint-to-vertex map (a-b+c, a+b-c, -a+b+c):
//   Ht = T (zeta + bz)                      vertex depths
//   if some Ht_i <= H0:
//     avg(Ht) <= H0: h_i = avg(Ht)          (flatten)
//     else: h[m] = H0; h[n] = max(H0, Ht[n] - (H0 - Ht[m])/2);
//           h[r] -= (h[m] - Ht[m]) + (h[n] - Ht[n])   (m, n, r as chosen
//           below; n == m when vertex 1 is the minimum, wetDry2D.c:115-124)
//     zeta_k = (h_k + h_k+1)/2 - bz_k
//   f_i = h_i > H0; no f_i: momentum zeroed; some f_i zero: vertex velocities
//   u = T q / h, the dry vertices' velocities (|T q| > VELZERO) summed and
//   shared among the wet ones, q_k = (h_k u_k + h_k+1 u_k+1)/2.
// ---------------------------------------------------------------------------
DGR_HD int pdop_adj(const double S[9], const double zv[3], double H0, double VZ, const double Sb[9], double Sin_b[9],
                    double zvb[3])
{
    const double bz[3] = {0.5 * (zv[0] + zv[1]), 0.5 * (zv[1] + zv[2]), 0.5 * (zv[2] + zv[0])};
    const double mh0 = S[0] + bz[0], mh1 = S[1] + bz[1], mh2 = S[2] + bz[2];
    const double Ht[3] = {mh0 - mh1 + mh2, mh0 + mh1 - mh2, -mh0 + mh1 + mh2};
    if (!(Ht[0] <= H0 || Ht[1] <= H0 || Ht[2] <= H0)) {
        for (int i = 0; i < 9; ++i) Sin_b[i] = Sb[i];
        return 0;
    }
    // forward replay of the depth correction
    const double avgH = (Ht[0] + Ht[1] + Ht[2]) / 3;
    const bool flat = avgH <= H0;
    double h[3] = {Ht[0], Ht[1], Ht[2]};
    int mi = 0, ni = 1, ri = 2;
    bool nact = false;   // the max(H0, .) took its second operand
    if (flat) {
        h[0] = h[1] = h[2] = avgH;
    } else {
        double mh = Ht[0];
        if (Ht[1] < mh) { mi = 1; mh = Ht[1]; }
        if (Ht[2] < mh) { mi = 2; mh = Ht[2]; }
        double nh = Ht[1];
        for (int i = 0; i < 3; ++i)
            if (Ht[i] < nh && i != mi) { ni = i; nh = Ht[i]; }
        for (int i = 2; i >= 0; --i)
            if (i != mi && i != ni) ri = i;
        h[mi] = H0;
        const double cand = nh - 0.5 * (h[mi] - mh);
        nact = cand > H0;
        h[ni] = nact ? cand : H0;
        h[ri] = h[ri] - (h[mi] - mh) - (h[ni] - nh);
    }
    // hb: adjoint of the corrected vertex depths h
    double hb[3] = {0, 0, 0};
    // momentum (both fields)
    const int f[3] = {h[0] > H0, h[1] > H0, h[2] > H0};
    const int npos = f[0] + f[1] + f[2];
    for (int q = 1; q < 3; ++q) {
        const double *Q = S + 3 * q;
        const double *ob = Sb + 3 * q;
        double *qb = Sin_b + 3 * q;
        if (npos == 0) {
            qb[0] = qb[1] = qb[2] = 0.0;
            continue;
        }
        if (npos == 3) {
            qb[0] = ob[0]; qb[1] = ob[1]; qb[2] = ob[2];
            continue;
        }
        const double qe[3] = {Q[0] - Q[1] + Q[2], Q[0] + Q[1] - Q[2], -Q[0] + Q[1] + Q[2]};
        double u[3], du = 0.0;
        for (int i = 0; i < 3; ++i) {
            u[i] = qe[i] / h[i];
            if (!f[i] && fabs(qe[i]) > VZ) du += u[i];
        }
        const double dun = du / npos;
        double qob[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) {
            qob[k] += 0.5 * ob[k];
            qob[(k + 1) % 3] += 0.5 * ob[k];
        }
        double ub[3], dunb = 0.0;
        for (int i = 0; i < 3; ++i) {
            const double un = f[i] ? u[i] + dun : 0.0;
            hb[i] += un * qob[i];
            const double unb = h[i] * qob[i];
            ub[i] = f[i] ? unb : 0.0;
            dunb += f[i] ? unb : 0.0;
        }
        const double dub = dunb / npos;
        double qeb[3];
        for (int i = 0; i < 3; ++i) {
            if (!f[i] && fabs(qe[i]) > VZ) ub[i] += dub;
            qeb[i] = ub[i] / h[i];
            hb[i] -= ub[i] * u[i] / h[i];
        }
        qb[0] = qeb[0] + qeb[1] - qeb[2];
        qb[1] = -qeb[0] + qeb[1] + qeb[2];
        qb[2] = qeb[0] - qeb[1] + qeb[2];
    }
    // zeta_k = (h_k + h_k+1)/2 - bz_k
    double bzb[3];
    for (int k = 0; k < 3; ++k) {
        hb[k] += 0.5 * Sb[k];
        hb[(k + 1) % 3] += 0.5 * Sb[k];
        bzb[k] = -Sb[k];
    }
    // the depth correction, reversed operation by operation
    double Htb[3];
    if (flat) {
        const double s = (hb[0] + hb[1] + hb[2]) / 3;
        Htb[0] = Htb[1] = Htb[2] = s;
    } else {
        double mhb = 0.0, nhb = 0.0;
        const double t3 = hb[ri];              // h[ri] -= (h[mi] - mh) + (h[ni] - nh)
        hb[mi] -= t3;
        hb[ni] -= t3;
        mhb += t3;
        nhb += t3;
        const double t2 = hb[ni];              // h[ni] = max(H0, nh - (h[mi] - mh)/2)
        hb[ni] = 0.0;
        if (nact) {
            nhb += t2;
            hb[mi] -= 0.5 * t2;
            mhb += 0.5 * t2;
        }
        hb[mi] = 0.0;                          // h[mi] = H0
        for (int i = 0; i < 3; ++i) Htb[i] = hb[i];
        Htb[mi] += mhb;
        Htb[ni] += nhb;
    }
    // Ht = T (zeta + bz)
    const double mhb3[3] = {Htb[0] + Htb[1] - Htb[2], -Htb[0] + Htb[1] + Htb[2], Htb[0] - Htb[1] + Htb[2]};
    for (int k = 0; k < 3; ++k) {
        Sin_b[k] = mhb3[k];
        bzb[k] += mhb3[k];
    }
    zvb[0] += 0.5 * (bzb[0] + bzb[2]);
    zvb[1] += 0.5 * (bzb[0] + bzb[1]);
    zvb[2] += 0.5 * (bzb[1] + bzb[2]);
    return 1;
}

#undef DGR_HD

}  // namespace dgr
