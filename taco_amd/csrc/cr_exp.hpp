// cr_exp.hpp — correctly rounded exp of a double in double-double arithmetic.
//
// The bath correlation function samples the spectral density J(w), whose
// Ohmic form carries exp(-w / w_c), and the KMS factors exp(-beta dw) and
// exp(beta w_Ny).  Their last bits reach Gamma and, through the FFT's
// rounding noise, L4 at the first time steps, so host and device must agree
// on them exactly.  The platform libm's exp is not correctly rounded; this
// routine returns the correctly rounded value, a platform-independent choice
// (the same convention as the phases in cr_sincos.hpp).
//
// Method: x = k ln2 + r with ln2 split into three doubles (products exact
// through fma), a table step r = j/64 + d with exp(j/64) stored as
// double-doubles, a degree-12 Taylor series of exp(d) for |d| <= 1/128 in
// double-double, and an exact scaling by 2^k (results in the subnormal range
// are rounded once, from the double-double, to the subnormal grid).  The
// result carries a relative error below ~2^-100 before its final rounding.
//
// Compiled with floating-point contraction off: the exact-transformation
// identities rely on separately rounded products and sums.
#pragma once

#include <cmath>

#include "cr_sincos.hpp"

namespace taco_cr {

constexpr int kExpJ = 23;                // exp(j/64) for j = -23..23 (|r| <= ln2/2 < 23/64)
constexpr int kExpTabN = 2 * kExpJ + 1;
constexpr int kExpDeg = 12;              // Taylor degree
// Table layout (doubles): [0, 2 kExpTabN) exp(j/64) as (hi, lo) for j + kExpJ;
// then 1/n! as (hi, lo) for n = 0..kExpDeg; then ln2 as three doubles.
constexpr int kExpTabInv = 2 * kExpTabN;
constexpr int kExpTabLn2 = kExpTabInv + 2 * (kExpDeg + 1);
constexpr int kExpTabSize = kExpTabLn2 + 3;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kExpMax = 0x1.62e42fefa39efp+9;    // above: the result overflows to +inf
constexpr double kExpMin = -0x1.74910d52d3052p+9;   // below (~ -745.1332): rounds to 0

TACO_HD inline double exp_cr(double x, const double* tab) {
    if (x != x) return x;
    if (x == 0.0) return 1.0;
    if (x > kExpMax) return HUGE_VAL;
    if (x < kExpMin) return 0.0;
    const double k = rint(x * kInvLn2);
    dd r;
    if (k == 0.0) {
        r = {x, 0.0};
    } else {
        const dd p0 = two_prod(k, tab[kExpTabLn2]);
        const double s0 = x - p0.hi;  // exact: k*L0 is within a factor 2 of x
        r = two_sum(s0, -p0.lo);
        r = dd_add(r, dd_neg(two_prod(k, tab[kExpTabLn2 + 1])));
        r = dd_add(r, dd_neg(two_prod(k, tab[kExpTabLn2 + 2])));
    }
    const double jr = rint(r.hi * 64.0);
    const int j = static_cast<int>(jr);
    dd d = r;
    if (j != 0) d = two_sum(r.hi - jr * 0.015625, r.lo);  // r.hi - j/64 is exact
    // exp(d) = sum_n d^n / n!, Horner from the highest term
    dd p = {tab[kExpTabInv + 2 * kExpDeg], tab[kExpTabInv + 2 * kExpDeg + 1]};
    for (int n = kExpDeg - 1; n >= 0; --n) p = dd_add(dd_mul(p, d), dd{tab[kExpTabInv + 2 * n], tab[kExpTabInv + 2 * n + 1]});
    if (j != 0) p = dd_mul(dd{tab[2 * (j + kExpJ)], tab[2 * (j + kExpJ) + 1]}, p);
    const int ki = static_cast<int>(k);
    // p.hi in [2^-1, 2^1): the result 2^k p is normal when k >= -1021
    if (ki >= -1021) return ldexp(p.hi, ki);
    // subnormal result: round 2^(k + 1074) p to an integer, then scale exactly
    const int sh = ki + 1074;
    const double hi = ldexp(p.hi, sh), lo = ldexp(p.lo, sh);
    double n = rint(hi);
    const double rem = (hi - n) + lo;  // hi - n is exact (|hi| < 2^53)
    if (rem > 0.5) n += 1.0;
    else if (rem < -0.5) n -= 1.0;
    else if (rem == 0.5 || rem == -0.5) {
        const double up = n + (rem > 0 ? 1.0 : -1.0);
        if (fmod(up, 2.0) == 0.0) n = up;
    }
    return ldexp(n, -1074);
}

}  // namespace taco_cr
