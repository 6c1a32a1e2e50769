// gfm_fastmath.hpp -- a plain-FP64 power function for the mixed-precision
// Riemann solver's Newton iteration.  It compiles for the host as well, so
// that a host build pins its results: every step is an IEEE operation or an
// explicit fma, with no contraction left to the compiler, so host and device
// produce the same bits.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define GFM_HD __host__ __device__
#else
#define GFM_HD
#endif

namespace gfm {

// x^y for finite x > 0 and |y log x| < 700, as exp(y log x):
//  log x = e ln2 + log m with x = m 2^e, m in [sqrt(1/2), sqrt(2)), and
//          log m = 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.1716, summed to
//          s^23 (truncation < 3e-18 relative);
//  exp t = 2^n exp(r), n = rint(t / ln2), |r| <= ln2/2, Taylor to r^13
//          (truncation < 5e-18 relative).
// Relative error about (2 + |y log x|) 1e-16.  x <= 0 or NaN gives NaN.
// Not inlined: its two dozen FP64 coefficients would otherwise be hoisted
// into registers of the calling kernel's loop.
GFM_HD inline __attribute__((noinline)) double pow_fast(double x, double y) {
#pragma clang fp contract(off)
    if (!(x > 0.0)) return x == 0.0 ? 0.0 : NAN;
    int e;
    double m = std::frexp(x, &e);                  // x = m 2^e, m in [0.5, 1)
    if (m < 0.70710678118654752440) { m = 2.0 * m; e -= 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double q = 1.0 / 23.0;
    q = std::fma(q, s2, 1.0 / 21.0);
    q = std::fma(q, s2, 1.0 / 19.0);
    q = std::fma(q, s2, 1.0 / 17.0);
    q = std::fma(q, s2, 1.0 / 15.0);
    q = std::fma(q, s2, 1.0 / 13.0);
    q = std::fma(q, s2, 1.0 / 11.0);
    q = std::fma(q, s2, 1.0 / 9.0);
    q = std::fma(q, s2, 1.0 / 7.0);
    q = std::fma(q, s2, 1.0 / 5.0);
    q = std::fma(q, s2, 1.0 / 3.0);
    const double s3 = s * s2;
    const double logm = std::fma(2.0 * s3, q, 2.0 * s);           // 2s + 2 s^3 q
    // ln 2 split so that e * LN2_HI is exact for |e| < 2^11
    const double LN2_HI = 6.93147180369123816490e-01, LN2_LO = 1.90821492927058770002e-10;
    const double de = (double)e;
    const double t = y * std::fma(de, LN2_LO, logm) + y * (de * LN2_HI);
    const double n = std::rint(t * 1.44269504088896340736);
    double r = std::fma(-n, LN2_HI, t);
    r = std::fma(-n, LN2_LO, r);
    double p = 1.0 / 6227020800.0;                                 // 1/13!
    p = std::fma(p, r, 1.0 / 479001600.0);
    p = std::fma(p, r, 1.0 / 39916800.0);
    p = std::fma(p, r, 1.0 / 3628800.0);
    p = std::fma(p, r, 1.0 / 362880.0);
    p = std::fma(p, r, 1.0 / 40320.0);
    p = std::fma(p, r, 1.0 / 5040.0);
    p = std::fma(p, r, 1.0 / 720.0);
    p = std::fma(p, r, 1.0 / 120.0);
    p = std::fma(p, r, 1.0 / 24.0);
    p = std::fma(p, r, 1.0 / 6.0);
    p = std::fma(p, r, 0.5);
    p = std::fma(p, r, 1.0);
    p = std::fma(p, r, 1.0);
    return std::ldexp(p, (int)n);
}

}  // namespace gfm
