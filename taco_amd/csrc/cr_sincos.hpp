// cr_sincos.hpp — correctly rounded sin/cos in double-double arithmetic.
//
// The F/C/R phases are exp(+-i*omega*(k*dt)) of a double argument.  Their
// last bits decide the FFT's rounding noise, which dominates L4(t) at the
// first time steps, so host and device must agree on them exactly.  The
// platform libm (glibc's sincos) is not correctly rounded; this routine
// returns the correctly rounded values, a unique platform-independent choice.
//
// Method: Cody–Waite reduction x = k*pi/2 + r with pi/2 split into three
// doubles (products exact through fma), a table step r = j/64 + d with
// sin/cos(j/64) stored as double-doubles, and Taylor series of sin(d),
// cos(d) for |d| <= 1/128 in double-double.  The result carries a relative
// error below ~2^-100 before its final rounding, so it is the correctly
// rounded value unless the exact one lies within 2^-100 of a rounding
// boundary (probability ~2^-47 per argument).  Valid for |x| < 2^30.
//
// Compiled with floating-point contraction off: the exact-transformation
// identities below rely on separately rounded products and sums.
#pragma once

#include <cmath>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define TACO_HD __host__ __device__
#else
#define TACO_HD
#endif

namespace taco_cr {

// pi/2 = P0 + P1 + P2 (+ O(2^-163)); each Pi = RN(pi/2 - sum of previous).
constexpr double kP0 = 0x1.921fb54442d18p+0;
constexpr double kP1 = 0x1.1a62633145c07p-54;
constexpr double kP2 = -0x1.f1976b7ed8fbcp-110;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr int kTabN = 52;  // sin(j/64), cos(j/64) for j = 0..51, as (hi, lo, hi, lo)
// 1/6, 1/24, 1/120 as double-doubles (hi = RN(1/n), lo = RN(1/n - hi)).
constexpr double kInv6Hi = 0x1.5555555555555p-3, kInv6Lo = 0x1.5555555555555p-57;
constexpr double kInv24Hi = 0x1.5555555555555p-5, kInv24Lo = 0x1.5555555555555p-59;
constexpr double kInv120Hi = 0x1.1111111111111p-7, kInv120Lo = 0x1.1111111111111p-63;

struct dd {
    double hi, lo;
};

TACO_HD inline dd two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
TACO_HD inline dd fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
TACO_HD inline dd two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}
TACO_HD inline dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
TACO_HD inline dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
TACO_HD inline dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo);
}
TACO_HD inline dd dd_div_d(dd a, double b) {
    const double q1 = a.hi / b;
    const dd p = two_prod(q1, b);
    const double r = ((a.hi - p.hi) - p.lo) + a.lo;
    return fast_two_sum(q1, r / b);
}

// tab: kTabN entries of (sin_hi, sin_lo, cos_hi, cos_lo) at j/64.
TACO_HD inline void sincos_cr(double x, double* s_out, double* c_out, const double* tab) {
    if (x == 0.0) {
        *s_out = x;
        *c_out = 1.0;
        return;
    }
    const double k = rint(x * kTwoOverPi);
    dd r;
    if (k == 0.0) {
        r = {x, 0.0};
    } else {
        const dd p0 = two_prod(k, kP0);
        const double s0 = x - p0.hi;  // exact: k*P0 is within a factor 2 of x
        r = two_sum(s0, -p0.lo);
        r = dd_add(r, dd_neg(two_prod(k, kP1)));
        r = dd_add(r, dd_neg(two_prod(k, kP2)));
    }
    const long long q = static_cast<long long>(k) & 3;
    const double jr = rint(r.hi * 64.0);
    const int j = static_cast<int>(jr);
    dd d = r;
    if (j != 0) d = two_sum(r.hi - jr * 0.015625, r.lo);  // r.hi - j/64 is exact
    const dd d2 = dd_mul(d, d);
    const dd d4 = dd_mul(d2, d2);
    const double x2 = d2.hi;
    const double x3 = x2 * x2 * x2;
    // sin d = d - d*v, v = d2/6 - d4/120 + d2^3/5040 - d2^4/362880
    dd v = dd_add(dd_mul(d2, dd{kInv6Hi, kInv6Lo}), dd_neg(dd_mul(d4, dd{kInv120Hi, kInv120Lo})));
    v = dd_add(v, dd{x3 * (1.0 / 5040.0) - x3 * x2 * (1.0 / 362880.0), 0.0});
    const dd sd = dd_add(d, dd_neg(dd_mul(d, v)));
    // cos d = 1 - w, w = d2/2 - d4/24 + d2^3/720 - d2^4/40320 + d2^5/3628800
    dd w = dd_add(dd{d2.hi * 0.5, d2.lo * 0.5}, dd_neg(dd_mul(d4, dd{kInv24Hi, kInv24Lo})));
    w = dd_add(w, dd{x3 * (1.0 / 720.0) - x3 * x2 * (1.0 / 40320.0) + x3 * x2 * x2 * (1.0 / 3628800.0), 0.0});
    const dd cd = dd_add(dd{1.0, 0.0}, dd_neg(w));
    dd sr = sd, cr = cd;
    if (j != 0) {
        const int aj = j < 0 ? -j : j;
        dd sc = {tab[4 * aj], tab[4 * aj + 1]};
        const dd cc = {tab[4 * aj + 2], tab[4 * aj + 3]};
        if (j < 0) sc = dd_neg(sc);
        sr = dd_add(dd_mul(sc, cd), dd_mul(cc, sd));
        cr = dd_add(dd_mul(cc, cd), dd_neg(dd_mul(sc, sd)));
    }
    double s, c;
    switch (q) {
        case 0: s = sr.hi; c = cr.hi; break;
        case 1: s = cr.hi; c = -sr.hi; break;
        case 2: s = -sr.hi; c = -cr.hi; break;
        default: s = -cr.hi; c = sr.hi; break;
    }
    *s_out = s;
    *c_out = c;
}

}  // namespace taco_cr
