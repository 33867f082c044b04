// The transform of float images into double pyramids and back
// (src/fixmath.h:165-186, :320-329): 1 + log1p(v) for v >= 0 and
// expm1(c - 1) for c >= 1.  The math library's log1p and expm1 carry
// extended-precision sequences of about 150 double operations each, which
// makes the pointwise conversion of a float image bound by double
// arithmetic rather than by memory.  These are written out from their
// series with a few dozen operations and stay within a few units in the
// last place of a double (the float results of the inverse round the same
// way but where a value lies within that error of a rounding boundary):
//   log(1 + v): u = 1 + v = m 2^k with m in [sqrt(1/2), sqrt(2)), f = m - 1
//     (exact), s = f / (2 + f), and log(1 + f) = 2 atanh(s)
//     = f - s (f - Q), Q = sum_{j>=1} 2 s^(2j) / (2j + 1), nine terms for
//     s^2 <= 0.0295; then k ln 2 added in two parts.  1 + v is exact for a
//     float v in [2^-29, 2^53]; below 2^-29, v - v^2 / 2 is within 2^-58;
//     above 2^53 the rounding of 1 + v moves the result by under 2^-53
//     relative.
//   e^x - 1: x = k ln 2 + r with |r| <= ln2 / 2 (Cody and Waite's two-part
//     reduction), e^r - 1 by its Taylor series to r^13 / 13!, then
//     2^k (e^r - 1) + (2^k - 1) with one rounding.
// Both compile for the host too, so the host tests can sweep them against
// the C library.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define ENBL_LOGEXP_FN __host__ __device__ __forceinline__
#else
#define ENBL_LOGEXP_FN inline
#endif

namespace enbl {

constexpr double kLn2Hi = 0x1.62e42fefa39efp-1;    // RN(ln 2)
constexpr double kLn2Lo = 0x1.abc9e3b39803fp-56;   // RN(ln 2 - kLn2Hi)
constexpr double kInvLn2 = 0x1.71547652b82fep+0;   // RN(1 / ln 2)

// log(1 + v) for v >= 0 (+inf gives +inf).
ENBL_LOGEXP_FN double log1p_nonneg(double v)
{
    const double u = 1.0 + v;
    int k;
    double m = frexp(u, &k);                       // u = m 2^k, m in [1/2, 1)
    const bool low = m < 0.70710678118654752440;
    m = low ? m + m : m;
    k = low ? k - 1 : k;
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double q = 2.0 / 19.0;
    q = fma(q, z, 2.0 / 17.0);
    q = fma(q, z, 2.0 / 15.0);
    q = fma(q, z, 2.0 / 13.0);
    q = fma(q, z, 2.0 / 11.0);
    q = fma(q, z, 2.0 / 9.0);
    q = fma(q, z, 2.0 / 7.0);
    q = fma(q, z, 2.0 / 5.0);
    q = fma(q, z, 2.0 / 3.0);
    q = q * z;
    const double t = fma(s, q - f, f);
    const double kd = (double)k;
    const double r = fma(kd, kLn2Hi, fma(kd, kLn2Lo, t));
    const double small = fma(-0.5 * v, v, v);
    return v < 0x1p-29 ? small : (v == (double)INFINITY ? v : r);
}

// e^x - 1 for x >= 0 (overflow gives +inf).
ENBL_LOGEXP_FN double expm1_nonneg(double x)
{
    if (x > 709.0) return (double)INFINITY;
    const double kd = rint(x * kInvLn2);
    const double r = fma(-kd, kLn2Lo, fma(-kd, kLn2Hi, x));
    double p = 1.0 / 6227020800.0;                 // 1 / 13!
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    const double em = fma(r * r, p, r);            // e^r - 1
    const int k = (int)kd;
    const double two_k = ldexp(1.0, k);
    return k == 0 ? em : fma(two_k, em, two_k - 1.0);
}

// The pyramid value of a float sample and back (src/fixmath.h:165-186,
// :320-329).
ENBL_LOGEXP_FN double float_to_log_pyramid(double v) { return v >= 0.0 ? 1.0 + log1p_nonneg(v) : 1.0 / (1.0 - v); }
ENBL_LOGEXP_FN double log_pyramid_to_float(double c) { return c >= 1.0 ? expm1_nonneg(c - 1.0) : 1.0 - 1.0 / c; }

}  // namespace enbl
