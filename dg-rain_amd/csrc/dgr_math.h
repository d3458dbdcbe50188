// dgr_math.h -- deterministic FP64 helpers for the gfx950 kernels.
//
// dgr_pow43 and dgr_sin use only IEEE-754 basic operations (+ - * / sqrt)
// and fused multiply-add, so they return the same bits on any IEEE machine.
//
//  dgr_pow43(x)  x^(4./3), the Manning term of the reference's friction
//                (2DCode/2DInnerProducts.c:318).  The reference calls glibc
//                pow(), which is not correctly rounded (about 0.08% of results
//                are 1 ulp off); no GPU routine can reproduce those bits.  This
//                one is accurate to ~1e-22 before the final rounding, i.e.
//                correctly rounded except in vanishingly rare cases.
//  dgr_sin(x)    sine for the tidal boundary forcing (no reference code).
//  dgr_log(x), dgr_log1p(x)   natural log of x >= 1 and of 1 + x, x >= 0, for
//                the Green-Ampt infiltration depth.
//  ga_cumulative Green-Ampt cumulative infiltration depth F(t).
//  DivExact / DivFast / DivRecip   division policies (below); DivFast gives
//                the bits of x / y with fewer instructions, DivRecip does not.
//
// Constants: see oracle/gen_constants.py for how each value was derived.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dgr {

// cubic least-squares seed for t^(-1/3) on [1,8) (max rel. error 1.8e-2)
constexpr double P43_C0 = 1.2316035297857837;
constexpr double P43_C1 = -0.29172246132464164;
constexpr double P43_C2 = 0.044484938428368009;
constexpr double P43_C3 = -0.0024458486365437917;
constexpr double THIRD = 0.33333333333333331;          // fl(1/3)
constexpr double DELTA43 = 7.4014868308343765e-17;     // 4/3 - fl(4/3) = 2^-52/3
constexpr double LN2 = 0.69314718055994529;

__host__ __device__ inline double pow2i(int k)  // 2^k, -1022 <= k <= 1023
{
    return __longlong_as_double((long long)((uint64_t)(k + 1023) << 52));
}

// x = m 2^e, e = 3q + r, t = m 2^r in [1,8): x^(4/3) = 2^(4q) t t^(1/3).
// t^(-1/3): seed + 3 Newton steps y += y(1 - t y^3)/3 (no division);
// t^(1/3) = t y^2, refined once in double-double; the exponent's representation
// error applied as the factor (1 - DELTA43 ln x).
// Branch-free: the special inputs (NaN, x <= 0, inf) run the arithmetic on
// whatever bits they have and their result is selected at the end, so the
// function adds no basic blocks to the kernels that inline it.
__device__ inline double dgr_pow43(double x)
{
    const bool den = x < 2.2250738585072014e-308;   // denormal (or not positive)
    const double xs = den ? x * 18014398509481984.0 : x;   // 2^54
    const uint64_t b = (uint64_t)__double_as_longlong(xs);
    int e = (int)((b >> 52) & 0x7ff) - (den ? 54 : 0);
    e -= 1023;
    const double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    const int q = (e + 3000) / 3 - 1000;   // floor(e / 3) for the exponent range
    const int r = e - 3 * q;
    const double t = m * (double)(1 << r);

    double y = fma(fma(fma(P43_C3, t, P43_C2), t, P43_C1), t, P43_C0);
#pragma unroll
    for (int it = 0; it < 3; ++it) {
        double w = y * y;
        w = w * y;
        const double h = fma(-t, w, 1.0);
        y = fma(y * h, THIRD, y);
    }
    const double c = (t * y) * y;
    const double a_hi = c * c;
    const double a_lo = fma(c, c, -a_hi);
    const double b_hi = a_hi * c;
    const double b_lo = fma(a_hi, c, -b_hi) + a_lo * c;
    const double d = (t - b_hi) - b_lo;
    const double dc = (d * (y * y)) * THIRD;
    const double p_hi = t * c;
    const double p_lo = fma(t, c, -p_hi) + t * dc;

    double mm = m;
    int ee = e;
    if (mm > 1.4142135623730951) { mm = mm * 0.5; ee += 1; }
    const double s = (mm - 1.0) / (mm + 1.0);
    const double s2 = s * s;
    const double lnm = 2.0 * s * fma(fma(fma(s2, 0.14285714285714285, 0.2), s2, 0.33333333333333331), s2, 1.0);
    const double lnx = fma((double)ee, LN2, lnm);
    const double eps = DELTA43 * lnx;

    const double res = p_hi + fma(-p_hi, eps, p_lo);
    const double sc = pow2i(2 * q);
    double out = (res * sc) * sc;
    // computed for every input (keeps the selects below from becoming branches)
    asm volatile("" : "+v"(out));
    out = (x > 1.7976931348623157e308) ? x : out;
    out = (x == 0.0) ? 0.0 : out;
    out = (x < 0.0) ? __longlong_as_double(0x7ff8000000000000LL) : out;
    return (x != x) ? x : out;
}

// ---------------------------------------------------------------------------
// Division policies.
//
// gfx950 divides doubles (correctly rounded) with
//   d = v_div_scale(y, y, x); r = v_rcp(d);
//   r = fma(r, fma(-d, r, 1), r) twice;
//   n = v_div_scale(x, y, x); q = n * r; e = fma(-d, q, n);
//   result = v_div_fixup(v_div_fmas(e, r, q), y, x).
// When no operand needs exponent scaling both div_scale steps return their
// first operand with the scale flag clear, div_fmas is the plain fma and
// div_fixup returns its first operand (up to the sign of a zero quotient,
// which it sets to sign(x) xor sign(y)).  Only q, e and the final fma then
// depend on x: DivFast keeps the rest per divisor (Den) and computes each
// quotient in 3 instructions, bit-identical to x / y.  It is exact for
//   divisors   |y| in [2^-320, 2^320]
//   numerators x = 0, or |x| in [2^-700, 2^320] (frexp exponent >= -699)
// (then |x/y| < 2^640 and nothing is scaled).  It checks this without
// branching, by folding the operands into three running bounds that ok()
// tests once: code evaluates a whole element (or edge) with DivFast and
// evaluates it again with DivExact when ok() fails.  A NaN operand gives NaN
// either way.  A zero quotient takes the sign of x*r, as div_fixup gives it
// (the final fma alone would round -0 + +0 to +0).
// DivRecip (reciprocal products, a few ulps off) is for values that only
// enter derivatives, never a forward result.
// ---------------------------------------------------------------------------
struct DivExact {
    struct Den {
        double y;
    };
    __device__ __forceinline__ Den den(double y) { return Den{y}; }
    __device__ __forceinline__ Den den_const(double y) { return Den{y}; }
    __device__ __forceinline__ double q(double x, const Den &d) { return x / d.y; }
};

struct DivFast {
    struct Den {
        double y, r;
    };
    int emin = 0;
    double amax = 0.0;
    double dmin = 0x1p320;

    __device__ __forceinline__ Den make(double y, double d)
    {
        double r = __builtin_amdgcn_rcp(d);
        double e = fma(-d, r, 1.0);
        r = fma(r, e, r);
        e = fma(-d, r, 1.0);
        r = fma(r, e, r);
        amax = fmax(amax, fabs(y));
        dmin = fmin(dmin, fabs(y));
        return Den{y, r};
    }
    __device__ __forceinline__ Den den(double y) { return make(y, y); }
    // a compile-time constant divisor must still go through v_rcp_f64
    __device__ __forceinline__ Den den_const(double y)
    {
        double d = y;
        asm volatile("" : "+v"(d));
        return make(y, d);
    }
    __device__ __forceinline__ double q(double x, const Den &d)
    {
        const int ex = __builtin_amdgcn_frexp_exp(x);
        emin = (ex < emin) ? ex : emin;
        amax = fmax(amax, fabs(x));
        const double qq = x * d.r;
        const double e = fma(-d.y, qq, x);
        return copysign(fma(e, d.r, qq), qq);
    }
    __device__ __forceinline__ bool ok() const { return emin >= -699 && amax <= 0x1p320 && dmin >= 0x1p-320; }
};

// reciprocal for values that only enter derivatives: the hardware estimate
// refined by two Newton steps (within about an ulp of 1/y, not correctly
// rounded); zero, infinite and NaN inputs give what 1/y gives.  Host builds
// divide.
__host__ __device__ __forceinline__ double recip_d(double y)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double r0 = __builtin_amdgcn_rcp(y);
    double r = fma(r0, fma(-y, r0, 1.0), r0);
    r = fma(r, fma(-y, r, 1.0), r);
    return __builtin_isfinite(r) ? r : r0;
#else
    return 1.0 / y;
#endif
}

// quotients as products with a reciprocal (within a few ulps of x / y): for
// values that only enter derivatives, never a forward result
struct DivRecip {
    struct Den {
        double r;
    };
    __device__ __forceinline__ Den den(double y) { return Den{recip_d(y)}; }
    __device__ __forceinline__ Den den_const(double y) { return Den{recip_d(y)}; }
    __device__ __forceinline__ double q(double x, const Den &d) { return x * d.r; }
};

// Cody-Waite reduction by pi/2 (33+33+53-bit split of pi/2) and Taylor
// polynomials through x^15 (sin) / x^16 (cos) on [-pi/4, pi/4].
__device__ inline double dgr_sin(double x)
{
    if (!(x - x == 0.0)) return x - x;
    const double TWO_OVER_PI = 0.63661977236758138;
    const double PIO2_1 = 1.5707963267341256e+00;
    const double PIO2_2 = 6.077100506303966e-11;
    const double PIO2_3 = 2.0222662487959506e-21;
    const double kd = rint(x * TWO_OVER_PI);
    double r = fma(-kd, PIO2_1, x);
    r = fma(-kd, PIO2_2, r);
    r = fma(-kd, PIO2_3, r);
    const long long k = (long long)kd;
    const double r2 = r * r;
    double sp = fma(r2, -7.6471637318198164e-13, 1.6059043836821613e-10);
    sp = fma(r2, sp, -2.5052108385441720e-08);
    sp = fma(r2, sp, 2.7557319223985893e-06);
    sp = fma(r2, sp, -1.9841269841269841e-04);
    sp = fma(r2, sp, 8.3333333333333332e-03);
    sp = fma(r2, sp, -1.6666666666666666e-01);
    const double sv = fma(r * r2, sp, r);
    double cp = fma(r2, 4.7794773323873853e-14, -1.1470745597729725e-11);
    cp = fma(r2, cp, 2.08767569878681e-09);
    cp = fma(r2, cp, -2.7557319223985888e-07);
    cp = fma(r2, cp, 2.4801587301587302e-05);
    cp = fma(r2, cp, -1.3888888888888889e-03);
    cp = fma(r2, cp, 4.1666666666666664e-02);
    cp = fma(r2, cp, -0.5);
    const double cv = fma(r2, cp, 1.0);
    switch ((int)(k & 3)) {
    case 0: return sv;
    case 1: return cv;
    case 2: return -sv;
    default: return -cv;
    }
}

// ln x for finite x >= 1: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m-1)/(m+1), by the odd series through s^21
// (truncation below 1e-18 relative), ln 2 split so that e * LN2_HI is exact.
// The oracle's orc_log is the same operations in the same order.
__device__ inline double dgr_log(double x)
{
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 21;
    p = p * s2 + 1.0 / 19;
    p = p * s2 + 1.0 / 17;
    p = p * s2 + 1.0 / 15;
    p = p * s2 + 1.0 / 13;
    p = p * s2 + 1.0 / 11;
    p = p * s2 + 1.0 / 9;
    p = p * s2 + 1.0 / 7;
    p = p * s2 + 1.0 / 5;
    p = p * s2 + 1.0 / 3;
    p = p * s2 + 1.0;
    const double lnm = 2.0 * s * p;
    return (double)e * 6.9314718036912382e-01 + ((double)e * 1.9082149292705877e-10 + lnm);
}

// ln(1 + x) for x >= 0 by Goldberg's correction (What Every Computer
// Scientist Should Know About Floating-Point Arithmetic, Theorem 4):
// u = 1 + x, ln(u) x / (u - 1), which keeps ln's relative accuracy where
// 1 + x rounds.  The oracle's orc_log1p is the same.
__device__ inline double dgr_log1p(double x)
{
    const double u = 1.0 + x;
    if (u == 1.0) return x;
    return dgr_log(u) * (x / (u - 1.0));
}

// Green-Ampt infiltration under ponding from t = 0 (the reference's kinematic
// model, KinField/Infiltration.c:8-87): the cumulative depth F solves
//   F - K t - P ln(1 + F/P) = 0,   P = psi * delta_theta.
// Newton from max(sqrt(2 P K t), K t) (below the root; the first step lands
// above it and the rest descend) until an iterate repeats, at most 40
// iterations; the stop depends only on the bits, so the oracle's
// orc_ga_cumulative, the same operations, stops at the same iterate.
__device__ inline double ga_cumulative(double K, double P, double t)
{
    if (!(t > 0)) return 0.0;
    const double Kt = K * t;
    double F = sqrt(2.0 * P * Kt);
    if (Kt > F) F = Kt;
    for (int i = 0; i < 40; ++i) {
        const double g = F - Kt - P * dgr_log1p(F / P);
        const double dg = F / (P + F);
        const double Fn = F - g / dg;
        if (Fn == F) break;
        F = Fn;
    }
    return F;
}

}  // namespace dgr
