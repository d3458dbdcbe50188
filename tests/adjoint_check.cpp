// adjoint_check.cpp -- host-side check of the reverse-mode pieces in
// dg-rain_amd/csrc/dgr_adjoint.h against central finite differences of the
// forward operations they differentiate (no GPU use).
//
// For a random point x, random output weights ybar and random directions d:
//     <ybar, (f(x + h d) - f(x - h d)) / 2h>  ==  <xbar, d>,  xbar = J^T ybar.
// Prints one line per check: name, max relative error over the directions.
// Exit status 0 if every check is within its tolerance.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../dg-rain_amd/csrc/dgr_device.h"
#include "../dg-rain_amd/csrc/dgr_adjoint.h"

using namespace dgr;

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

struct GeoArr {
    double v[GX_COUNT];
    double operator()(int i) const { return v[i]; }
};

// forward of the element RHS (2DInnerProducts.c:14-566, compute2DL.c:344-409)
// in plain double arithmetic over an arbitrary geometry vector
static void rhs_eval(const GeoArr &G, const double W[9], const double Fh[9], int wd_self, double g, double R[9])
{
    const double nx[3] = {G(GX_NX), G(GX_NX + 1), G(GX_NX + 2)}, ny[3] = {G(GX_NY), G(GX_NY + 1), G(GX_NY + 2)};
    const Rt0 T = rt0_map(nx[0], G(GX_M2), ny[0], nx[1], G(GX_M2 + 1), ny[1], nx[2], G(GX_M2 + 2), ny[2], G(GX_DEN));
    double SI[9] = {0};
    for (int f = 0; f < 3; ++f) {
        double a = 0, b = 0, c = 0;
        for (int k = 0; k < 3; ++k) { a += T.A[k] * Fh[f * 3 + k]; b += T.B[k] * Fh[f * 3 + k]; c += T.C[k] * Fh[f * 3 + k]; }
        for (int ed = 0; ed < 3; ++ed)
            for (int n = 0; n < 2; ++n) {
                const double xq = G(GX_XE + 2 * ed + n), yq = G(GX_YE + 2 * ed + n);
                const double Fn = (a + b * xq) * nx[ed] + (c + b * yq) * ny[ed];
                for (int bb = 0; bb < 3; ++bb) SI[f * 3 + bb] += G(GX_EJ + ed) * Fn * G(GX_PS + 6 * bb + 2 * ed + n);
            }
    }
    const double lg = wd_self == 0 ? 0.0 : g, weight = 1.0 / 6;
    double tau[3], Fx[3][3], Fy[3][3];
    for (int k = 0; k < 3; ++k) {
        const double mz = G(GX_MZ + k), H = W[k] + mz, qx = W[3 + k], qy = W[6 + k];
        const double u = qx / H, v = qy / H;
        tau[k] = G(GX_GNN + k) * std::sqrt(u * u + v * v) / (H * std::cbrt(H));
        Fx[0][k] = qx; Fx[1][k] = qx * qx / H + 0.5 * lg * (H * H - mz * mz); Fx[2][k] = qx * qy / H;
        Fy[0][k] = qy; Fy[1][k] = qx * qy / H; Fy[2][k] = qy * qy / H + 0.5 * lg * (H * H - mz * mz);
    }
    for (int f = 0; f < 3; ++f) {
        double r[3];
        for (int k = 0; k < 3; ++k) r[k] = Fx[f][k] * nx[k] + Fy[f][k] * ny[k];
        double va = 0, vb = 0, vc = 0;
        for (int k = 0; k < 3; ++k) { va += T.A[k] * r[k]; vb += T.B[k] * r[k]; vc += T.C[k] * r[k]; }
        for (int bb = 0; bb < 3; ++bb) {
            const double VF = G(GX_AJ) * (0.5 * va * G(GX_GP + 2 * bb) + 0.5 * vc * G(GX_GP + 2 * bb + 1) +
                                          vb * weight * G(GX_GS + bb));
            double e = -SI[f * 3 + bb] + VF;
            if (f == 1) e += (wd_self ? G(GX_BGX) * W[bb] : 0.0) - G(GX_AW) * tau[bb] * W[3 + bb];
            if (f == 2) e += (wd_self ? G(GX_BGY) * W[bb] : 0.0) - G(GX_AW) * tau[bb] * W[6 + bb];
            R[f * 3 + bb] = G(GX_MI) * e;
        }
    }
}

// limiter element pass (SlopeLimiter2D.c:157-343) in plain double arithmetic
static void limit_eval(const double S[9], const double Vmin[3][3], const double Vmax[3][3], const double zv[3],
                       double H0, double out[9])
{
    int modify = 0;
    for (int i = 0; i < 9; ++i) out[i] = S[i];
    double res[9];
    for (int iv = 0; iv < 3; ++iv) {
        const double w1 = S[3 * iv], w2 = S[3 * iv + 1], w3 = S[3 * iv + 2];
        double V[3] = {w1 - w2 + w3, w1 + w2 - w3, -w1 + w2 + w3};
        const double avg = (w1 + w2 + w3) / 3;
        if (iv == 0 && V[0] + zv[0] > H0 && V[1] + zv[1] > H0 && V[2] + zv[2] > H0 && Vmin[0][0] + zv[0] > H0 &&
            Vmin[0][1] + zv[1] > H0 && Vmin[0][2] + zv[2] > H0)
            modify = 1;
        for (int k = 0; k < 3; ++k) V[k] = std::fmax(V[k] < Vmax[iv][k] ? V[k] : Vmax[iv][k], Vmin[iv][k]);
        for (int ll = 0; ll < 3; ++ll) {
            double sumdiff = ((V[0] + V[1] + V[2]) / 3 - avg) * 3;
            const int sd = sumdiff < 0 ? -1 : (sumdiff == 0 ? 0 : 1);
            const double diff[3] = {(V[0] - avg) * sd, (V[1] - avg) * sd, (V[2] - avg) * sd};
            double kdp = (diff[0] > 0) + (diff[1] > 0) + (diff[2] > 0);
            for (int k = 0; k < 3; ++k) {
                const double dv = std::fmax(1.0, kdp);
                double red = 0;
                if (diff[k] > 0) { red = sumdiff * sd / dv; kdp -= 1; }
                const double redmax = sd > 0 ? V[k] - Vmin[iv][k] : Vmax[iv][k] - V[k];
                red = red < redmax ? red : redmax;
                sumdiff -= red * sd;
                V[k] -= red * sd;
            }
        }
        if (iv == 0 && (V[0] + zv[0] < H0 || V[1] + zv[1] < H0 || V[2] + zv[2] < H0)) modify = 0;
        res[3 * iv] = (V[0] + V[1]) / 2;
        res[3 * iv + 1] = (V[1] + V[2]) / 2;
        res[3 * iv + 2] = (V[2] + V[0]) / 2;
    }
    if (modify)
        for (int i = 0; i < 9; ++i) out[i] = res[i];
}

// PDop2D on one element (wetDry2D.c:66-256) in plain double arithmetic
static void pdop_eval(const double S[9], const double zv[3], double H0, double VZ, double out[9])
{
    for (int i = 0; i < 9; ++i) out[i] = S[i];
    const double mz[3] = {0.5 * (zv[0] + zv[1]), 0.5 * (zv[1] + zv[2]), 0.5 * (zv[2] + zv[0])};
    const double m0 = S[0] + mz[0], m1 = S[1] + mz[1], m2 = S[2] + mz[2];
    double Ht[3] = {m0 - m1 + m2, m0 + m1 - m2, -m0 + m1 + m2};
    const double avgH = (Ht[0] + Ht[1] + Ht[2]) / 3;
    if (Ht[0] <= H0 || Ht[1] <= H0 || Ht[2] <= H0) {
        if (avgH <= H0) {
            for (int i = 0; i < 3; ++i) Ht[i] = avgH;
        } else {
            int mi = 0;
            double mh = Ht[0];
            for (int i = 1; i < 3; ++i)
                if (Ht[i] < mh) { mi = i; mh = Ht[i]; }
            int ni = 1;
            double nh = Ht[1];
            for (int i = 0; i < 3; ++i)
                if (Ht[i] < nh && i != mi) { ni = i; nh = Ht[i]; }
            int ri = 2;
            for (int i = 0; i < 3; ++i)
                if (i != mi && i != ni) { ri = i; break; }
            Ht[mi] = H0;
            Ht[ni] = std::fmax(H0, nh - 0.5 * (Ht[mi] - mh));
            Ht[ri] = Ht[ri] - (Ht[mi] - mh) - (Ht[ni] - nh);
        }
        for (int k = 0; k < 3; ++k) out[k] = 0.5 * (Ht[k] + Ht[(k + 1) % 3]) - mz[k];
    }
    int f[3], npos = 0;
    for (int i = 0; i < 3; ++i) { f[i] = Ht[i] > H0; npos += f[i]; }
    if (npos == 3) return;
    for (int q = 1; q < 3; ++q) {
        const double *Q = S + 3 * q;
        if (npos == 0) { out[3 * q] = out[3 * q + 1] = out[3 * q + 2] = 0; continue; }
        const double qe[3] = {Q[0] - Q[1] + Q[2], Q[0] + Q[1] - Q[2], -Q[0] + Q[1] + Q[2]};
        double u[3], du = 0;
        for (int i = 0; i < 3; ++i) {
            u[i] = qe[i] / Ht[i];
            du += u[i] * (1 - f[i]) * (std::fabs(qe[i]) > VZ);
        }
        double qo[3];
        for (int i = 0; i < 3; ++i) qo[i] = Ht[i] * (f[i] * (u[i] + du / npos));
        for (int k = 0; k < 3; ++k) out[3 * q + k] = 0.5 * (qo[k] + qo[(k + 1) % 3]);
    }
}

static int report(const char *name, double err, double tol)
{
    printf("%-28s max rel err %.3e (tol %.0e) %s\n", name, err, tol, err <= tol ? "ok" : "FAIL");
    return err <= tol ? 0 : 1;
}


static int check_roe(int trials)
{
    double worst = 0;
    const double g = 9.81;
    for (int t = 0; t < trials; ++t) {
        RoeIn a;
        a.zb = U(1, 8);
        a.zi = U(-0.5, 2); a.ze = a.zi + U(-0.3, 0.3);
        a.qxi = U(-3, 3); a.qxe = U(-3, 3); a.qyi = U(-3, 3); a.qye = U(-3, 3);
        const double th = U(0, 6.283185307179586), nx = std::cos(th), ny = std::sin(th);
        const double lg = (t % 3 == 0) ? 0.0 : g;
        const double Fb[3] = {U(-1, 1), U(-1, 1), U(-1, 1)}, lamb = (t % 2) ? U(-1, 1) : 0.0;
        RoeIn ib = {0, 0, 0, 0, 0, 0, 0};
        roe_adj(a, nx, ny, lg, g, Fb, lamb, ib);
        for (int dir = 0; dir < 4; ++dir) {
            double d[7], x[7] = {a.zi, a.ze, a.qxi, a.qxe, a.qyi, a.qye, a.zb};
            for (double &v : d) v = U(-1, 1);
            const double h = 1e-6;
            double yp = 0, ym = 0;
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                RoeIn b;
                double *pb = &b.zi;
                for (int i = 0; i < 7; ++i) pb[i] = x[i] + sgn * h * d[i];
                double F[3];
                const double lam = roe_eval(F, b, nx, ny, lg, g);
                const double y = Fb[0] * F[0] + Fb[1] * F[1] + Fb[2] * F[2] + lamb * lam;
                (sgn < 0 ? ym : yp) = y;
            }
            const double fd = (yp - ym) / (2 * h);
            const double *pib = &ib.zi;
            double ad = 0, mag = 0;
            for (int i = 0; i < 7; ++i) { ad += pib[i] * d[i]; mag += std::fabs(pib[i] * d[i]); }
            const double e = std::fabs(ad - fd) / (mag + 1e-300);
            worst = e > worst ? e : worst;
        }
    }
    return report("roe_adj", worst, 1e-6);
}

static int check_rhs(int trials)
{
    double worst = 0;
    const double g = 9.81;
    for (int t = 0; t < trials; ++t) {
        GeoArr G;
        for (double &v : G.v) v = U(-1, 1);
        for (int k = 0; k < 3; ++k) {
            G.v[GX_MZ + k] = U(2, 8);
            G.v[GX_GNN + k] = U(0.002, 0.01);
            G.v[GX_EJ + k] = U(5, 20);
        }
        G.v[GX_DEN] = U(1, 2);
        G.v[GX_MI] = U(0.1, 1);
        G.v[GX_AJ] = U(10, 50);
        G.v[GX_AW] = G.v[GX_AJ] / 6;
        const int wd = (t % 4 == 0) ? 0 : 1;
        double W[9], Fh[9], Rb[9];
        for (int k = 0; k < 3; ++k) { W[k] = U(-0.5, 0.5); W[3 + k] = U(-2, 2); W[6 + k] = U(-2, 2); }
        for (double &v : Fh) v = U(-3, 3);
        for (double &v : Rb) v = U(-1, 1);
        double Wb[9] = {0}, Fhb[9] = {0}, mzb[3] = {0}, gnnb[3] = {0}, bgxb = 0, bgyb = 0;
        element_rhs_adj(G, W, wd, g, Rb, Wb, Fhb, mzb, gnnb, bgxb, bgyb);
        for (int dir = 0; dir < 4; ++dir) {
            double dW[9], dF[9], dmz[3], dgn[3], dbx = U(-1, 1), dby = U(-1, 1);
            for (double &v : dW) v = U(-1, 1);
            for (double &v : dF) v = U(-1, 1);
            for (double &v : dmz) v = U(-1, 1);
            for (double &v : dgn) v = U(-0.001, 0.001);
            const double h = 1e-6;
            double y[2];
            for (int s = 0; s < 2; ++s) {
                const double sg = s ? 1.0 : -1.0;
                GeoArr G2 = G;
                double W2[9], F2[9], R[9];
                for (int i = 0; i < 9; ++i) { W2[i] = W[i] + sg * h * dW[i]; F2[i] = Fh[i] + sg * h * dF[i]; }
                for (int k = 0; k < 3; ++k) {
                    G2.v[GX_MZ + k] += sg * h * dmz[k];
                    G2.v[GX_GNN + k] += sg * h * dgn[k];
                }
                G2.v[GX_BGX] += sg * h * dbx;
                G2.v[GX_BGY] += sg * h * dby;
                rhs_eval(G2, W2, F2, wd, g, R);
                y[s] = 0;
                for (int i = 0; i < 9; ++i) y[s] += Rb[i] * R[i];
            }
            const double fd = (y[1] - y[0]) / (2 * h);
            double ad = bgxb * dbx + bgyb * dby, mag = std::fabs(bgxb * dbx) + std::fabs(bgyb * dby);
            for (int i = 0; i < 9; ++i) { ad += Wb[i] * dW[i] + Fhb[i] * dF[i]; mag += std::fabs(Wb[i] * dW[i]) + std::fabs(Fhb[i] * dF[i]); }
            for (int k = 0; k < 3; ++k) { ad += mzb[k] * dmz[k] + gnnb[k] * dgn[k]; mag += std::fabs(mzb[k] * dmz[k]) + std::fabs(gnnb[k] * dgn[k]); }
            const double e = std::fabs(ad - fd) / (mag + 1e-300);
            worst = e > worst ? e : worst;
        }
    }
    return report("element_rhs_adj", worst, 1e-6);
}

static int check_limit(int trials)
{
    double worst = 0;
    int active = 0;
    for (int t = 0; t < trials; ++t) {
        double S[9], Vmin[3][3], Vmax[3][3], zv[3], Sb[9];
        for (int k = 0; k < 3; ++k) zv[k] = U(3, 6);
        for (int f = 0; f < 3; ++f) {
            const double base = f == 0 ? U(-0.2, 0.2) : U(-1, 1);
            for (int k = 0; k < 3; ++k) S[3 * f + k] = base + U(-0.3, 0.3);
            for (int k = 0; k < 3; ++k) {
                Vmin[f][k] = base + U(-0.35, -0.02);
                Vmax[f][k] = base + U(0.02, 0.35);
            }
        }
        for (double &v : Sb) v = U(-1, 1);
        double Sin_b[9] = {0}, Vminb[3][3] = {{0}}, Vmaxb[3][3] = {{0}};
        active += limit_adj(S, Vmin, Vmax, zv, 1e-3, Sb, Sin_b, Vminb, Vmaxb);
        for (int dir = 0; dir < 4; ++dir) {
            double dS[9], dmin[3][3], dmax[3][3];
            for (double &v : dS) v = U(-1, 1);
            for (int f = 0; f < 3; ++f)
                for (int k = 0; k < 3; ++k) { dmin[f][k] = U(-1, 1); dmax[f][k] = U(-1, 1); }
            const double h = 1e-9;   // small enough to stay on one branch
            double y[2];
            for (int s = 0; s < 2; ++s) {
                const double sg = s ? 1.0 : -1.0;
                double S2[9], mn[3][3], mx[3][3], out[9];
                for (int i = 0; i < 9; ++i) S2[i] = S[i] + sg * h * dS[i];
                for (int f = 0; f < 3; ++f)
                    for (int k = 0; k < 3; ++k) { mn[f][k] = Vmin[f][k] + sg * h * dmin[f][k]; mx[f][k] = Vmax[f][k] + sg * h * dmax[f][k]; }
                limit_eval(S2, mn, mx, zv, 1e-3, out);
                y[s] = 0;
                for (int i = 0; i < 9; ++i) y[s] += Sb[i] * out[i];
            }
            const double fd = (y[1] - y[0]) / (2 * h);
            double ad = 0, mag = 0;
            for (int i = 0; i < 9; ++i) { ad += Sin_b[i] * dS[i]; mag += std::fabs(Sin_b[i] * dS[i]); }
            for (int f = 0; f < 3; ++f)
                for (int k = 0; k < 3; ++k) {
                    ad += Vminb[f][k] * dmin[f][k] + Vmaxb[f][k] * dmax[f][k];
                    mag += std::fabs(Vminb[f][k] * dmin[f][k]) + std::fabs(Vmaxb[f][k] * dmax[f][k]);
                }
            const double e = std::fabs(ad - fd) / (mag + 1e-300);
            if (e > 1e-5 && getenv("ADJ_VERBOSE")) printf("limit trial %d dir %d ad %.10e fd %.10e\n", t, dir, ad, fd);
            worst = e > worst ? e : worst;
        }
    }
    printf("limit_adj: %d of %d elements limited\n", active, trials);
    return report("limit_adj", worst, 1e-5);
}

// the packed form of a field pass's decisions (lim_dec_pack) gives back
// every decision, and the reverse from the unpacked form equals, bit for
// bit, the reverse from the pass's own record; random passes with clipped
// and unclipped vertices, both signs of the redistribution and ties
static int check_lim_dec(int trials)
{
    int bad = 0, clipped = 0;
    for (int t = 0; t < trials; ++t) {
        double S[9], Vmin[3], Vmax[3], Sb[9];
        const int iv = t % 3;
        const double base = U(-1, 1);
        for (double &v : S) v = base + U(-0.3, 0.3);
        for (int k = 0; k < 3; ++k) {
            Vmin[k] = base + U(-0.35, -0.02);
            Vmax[k] = base + U(0.02, 0.35);
        }
        if (t % 7 == 0) S[3 * iv] = S[3 * iv + 1] = S[3 * iv + 2] = base;   // sumdiff 0 at once
        for (double &v : Sb) v = U(-1, 1);
        LimFieldDec d;
        limit_field_forward(S, iv, Vmin, Vmax, d);
        const LimFieldDec u = lim_dec_unpack(lim_dec_pack(d));
        bool same = d.pos == u.pos && d.use0 == u.use0;
        for (int k = 0; k < 3; ++k) {
            same = same && d.clampsel[k] == u.clampsel[k] && d.sd[k] == u.sd[k];
            clipped += d.clampsel[k] != 0;
        }
        double a[9] = {0}, b[9] = {0}, amn[3] = {0}, amx[3] = {0}, bmn[3] = {0}, bmx[3] = {0};
        limit_field_reverse(d, iv, Sb, a, amn, amx);
        limit_field_reverse(u, iv, Sb, b, bmn, bmx);
        for (int i = 0; i < 9; ++i) same = same && std::memcmp(&a[i], &b[i], 8) == 0;
        for (int k = 0; k < 3; ++k)
            same = same && std::memcmp(&amn[k], &bmn[k], 8) == 0 && std::memcmp(&amx[k], &bmx[k], 8) == 0;
        bad += !same;
    }
    printf("lim_dec_pack: %d of %d passes differ after packing (%d clipped vertices)\n", bad, trials, clipped);
    return bad ? 1 : 0;
}

static int check_pdop(int trials)
{
    const double H0 = 1e-3, VZ = 1e-2;
    double worst = 0;
    int acted = 0, flat = 0, min1 = 0, zeroed = 0, moved = 0;
    for (int t = 0; t < trials; ++t) {
        // vertex depths around H0 (some dry), the midpoint state that has them
        double zv[3], Hv[3], S[9], Sb[9];
        for (int k = 0; k < 3; ++k) zv[k] = U(0.01, 0.05);
        const int kind = t % 4;
        for (int k = 0; k < 3; ++k)
            Hv[k] = kind == 0 ? U(-2e-3, 1.5e-3)                   // mostly flattened
                  : kind == 1 ? U(-3e-3, 0.05)                     // lift the minimum
                  : kind == 2 ? (k == 1 ? U(-3e-3, 8e-4) : U(4e-3, 0.05))   // vertex 1 the minimum
                              : U(-1e-3, 0.02);
        for (int k = 0; k < 3; ++k) S[k] = 0.5 * (Hv[k] + Hv[(k + 1) % 3]) - 0.5 * (zv[k] + zv[(k + 1) % 3]);
        for (int k = 3; k < 9; ++k) S[k] = U(-0.03, 0.03);
        for (double &v : Sb) v = U(-1, 1);
        double Sin_b[9], zvb[3] = {0, 0, 0};
        const int a = pdop_adj(S, zv, H0, VZ, Sb, Sin_b, zvb);
        acted += a;
        if (a) {
            const double avg = (Hv[0] + Hv[1] + Hv[2]) / 3;
            flat += avg <= H0;
            min1 += avg > H0 && Hv[1] < Hv[0] && Hv[1] <= Hv[2];
            double o[9];
            pdop_eval(S, zv, H0, VZ, o);
            zeroed += o[3] == 0 && o[4] == 0 && o[5] == 0;
            moved += !(o[3] == 0 && o[4] == 0 && o[5] == 0) && (o[3] != S[3] || o[4] != S[4] || o[5] != S[5]);
        }
        for (int dir = 0; dir < 4; ++dir) {
            double dS[9], dz[3];
            for (double &v : dS) v = U(-1, 1);
            for (double &v : dz) v = U(-1, 1);
            const double h = 1e-9;
            double y[2];
            for (int s = 0; s < 2; ++s) {
                const double sg = s ? 1.0 : -1.0;
                double S2[9], z2[3], out[9];
                for (int i = 0; i < 9; ++i) S2[i] = S[i] + sg * h * dS[i];
                for (int k = 0; k < 3; ++k) z2[k] = zv[k] + sg * h * dz[k];
                pdop_eval(S2, z2, H0, VZ, out);
                y[s] = 0;
                for (int i = 0; i < 9; ++i) y[s] += Sb[i] * out[i];
            }
            const double fd = (y[1] - y[0]) / (2 * h);
            double ad = 0, mag = 0;
            for (int i = 0; i < 9; ++i) { ad += Sin_b[i] * dS[i]; mag += std::fabs(Sin_b[i] * dS[i]); }
            for (int k = 0; k < 3; ++k) { ad += zvb[k] * dz[k]; mag += std::fabs(zvb[k] * dz[k]); }
            const double e = std::fabs(ad - fd) / (mag + 1e-300);
            if (e > 1e-5 && getenv("ADJ_VERBOSE")) printf("pdop trial %d dir %d ad %.10e fd %.10e\n", t, dir, ad, fd);
            worst = e > worst ? e : worst;
        }
    }
    printf("pdop_adj: %d of %d elements changed (%d flattened, %d with vertex 1 the minimum, %d momentum zeroed, "
           "%d momentum redistributed)\n", acted, trials, flat, min1, zeroed, moved);
    int bad = report("pdop_adj", worst, 1e-6);
    if (flat == 0 || min1 == 0 || zeroed == 0 || moved == 0 || acted - flat == 0) {
        printf("pdop_adj: a branch was not exercised\n");
        bad = 1;
    }
    return bad;
}

int main()
{
    int bad = 0;
    bad += check_roe(2000);
    bad += check_rhs(500);
    bad += check_limit(2000);
    bad += check_lim_dec(20000);
    bad += check_pdop(4000);
    return bad ? 1 : 0;
}
