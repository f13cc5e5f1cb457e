// hfv_pow.hpp -- pow(x, y) in double-double arithmetic, for the pressures
// of the two-rarefaction guess (positive x, y = (gamma-1)/(2 gamma) or its
// inverse), without a call to the library pow on any argument.
//
// Method (standard table-based log and exp in double-double arithmetic):
//   x = 2^k m, m in [0.75, 1.5); c = RN(1/m_i) for m's interval i of width
//   1/64; r = m c - 1 is carried exactly as a double-double (m c by a
//   two-product, then - 1 exactly, since m c lies in [0.5, 2]);
//   log x = k ln2 + (-log c) + log1p(r), |r| < 2^-6.5, log1p by its Taylor
//   series to r^9, all sums as double-doubles: absolute error ~2^-66.
//   w = y log x as a double-double; w = n ln2/64 + r (Cody-Waite, n*ln2/64
//   split so n*LN2_64_HI is exact), |r| <= ln2/128; exp(r) - 1 by its
//   Taylor series to r^6; 2^(n/64) = 2^(n>>6) * 2^(j/64) with 2^(j/64) a
//   double-double table entry.
// For positive finite x (subnormal included) with a normal result, the
// value before its last rounding is within ~2^-64 relative of x^y, so it is
// the correctly rounded x^y except when x^y lies within ~2^-11 ulp of a
// rounding boundary.  Results in the subnormal range are rounded twice
// (within 1 ulp), and where y log x lies within ~1e-13 below
// ln(DBL_MAX) the result may be DBL_MAX where x^y rounds to infinity.
// Zero, negative, infinite and NaN arguments follow C99 F.9.4.4 (pow),
// evaluated by selects so that no lane branches.
// Tables: hfv_pow_tables.hpp (gen_pow_tables.py).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "hfv_pow_tables.hpp"

#ifndef HFV_HD
#define HFV_HD __host__ __device__
#endif

// Host builds of the reference's own C API (hfv_physics.cpp) define
// HFV_HOST_LIBM_POW: there every power is the host C library's pow(), the
// function the reference calls, so those entry points return the
// reference's bits.  Device code always takes the double-double path.
#if defined(HFV_HOST_LIBM_POW) && !defined(__HIP_DEVICE_COMPILE__)
#define HFV_LIBM_POW 1
#endif

namespace hfv {

struct DD {
    double hi, lo;
};

// Knuth's two-sum: a + b = s + e exactly
HFV_HD __forceinline__ DD two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return DD{s, e};
}

// a + b = s + e exactly when |a| >= |b| (or a == 0)
HFV_HD __forceinline__ DD fast_two_sum(double a, double b)
{
    const double s = a + b;
    return DD{s, b - (s - a)};
}

HFV_HD __forceinline__ const double* pow_log_tab()
{
#ifdef __HIP_DEVICE_COMPILE__
    return POW_LOG_TAB_D;
#else
    return POW_LOG_TAB_H;
#endif
}

HFV_HD __forceinline__ const double* pow_exp_tab()
{
#ifdef __HIP_DEVICE_COMPILE__
    return POW_EXP_TAB_D;
#else
    return POW_EXP_TAB_H;
#endif
}

// log x as a double-double, x positive finite normal; lt: the log table
// (POW_LOG_TAB_*, or a copy of it)
HFV_HD __forceinline__ DD log_dd(double x, const double* lt = pow_log_tab())
{
    int k;
    double m = frexp(x, &k);  // m in [0.5, 1)
    if (m < 0.75) {
        m *= 2.0;
        k -= 1;
    }
    const int i = (int)(m * 64.0) - 48;  // m in [0.75 + i/64, 0.75 + (i+1)/64)
    const double* t = lt + 3 * i;
    const double c = t[0];
    // r = m c - 1 exactly: ph + pl = m c, ph - 1 exact (ph in [0.5, 2])
    const double ph = m * c;
    const double pl = fma(m, c, -ph);
    const double rh = ph - 1.0;
    // log1p(rh + pl) = rh + [pl (1 - rh) - rh^2 (1/2 - rh/3 + ... - rh^7/9)]
    double q = -1.0 / 9.0;
    q = fma(q, rh, 1.0 / 8.0);
    q = fma(q, rh, -1.0 / 7.0);
    q = fma(q, rh, 1.0 / 6.0);
    q = fma(q, rh, -1.0 / 5.0);
    q = fma(q, rh, 1.0 / 4.0);
    q = fma(q, rh, -1.0 / 3.0);
    q = fma(q, rh, 1.0 / 2.0);
    const double tail = fma(pl, -rh, pl) - (rh * rh) * q;
    // k ln2 + (-log c) + rh + (low parts)
    const double kd = (double)k;
    const DD s1 = two_sum(kd * POW_LN2_HI, t[1]);
    const DD s2 = two_sum(s1.hi, rh);
    const double low = ((s1.lo + s2.lo) + (kd * POW_LN2_LO + t[2])) + tail;
    return fast_two_sum(s2.hi, low);
}

// exp(w.hi + w.lo) for |w.hi| < 700; et: the exp table (POW_EXP_TAB_*, or a copy)
HFV_HD __forceinline__ double exp_dd(DD w, const double* et = pow_exp_tab())
{
    const double nd = rint(w.hi * POW_64_OVER_LN2);
    const int n = (int)nd;
    const int j = n & 63;
    const int e = n >> 6;  // floor(n / 64)
    // r = w - n ln2/64 as a double-double: the high difference is exact, and
    // n*LN2_64_LO is not small next to it (up to ~2^-30), so renormalise
    const DD r = two_sum(fma(-nd, POW_LN2_64_HI, w.hi), fma(-nd, POW_LN2_64_LO, w.lo));
    const double rh = r.hi, rl = r.lo;
    // exp(rh + rl) - 1 = rh + [rl (1 + rh) + rh^2 (1/2 + rh/6 + ... + rh^4/720)]
    double q = 1.0 / 720.0;
    q = fma(q, rh, 1.0 / 120.0);
    q = fma(q, rh, 1.0 / 24.0);
    q = fma(q, rh, 1.0 / 6.0);
    q = fma(q, rh, 1.0 / 2.0);
    const double b = fma(rl, rh, rl) + (rh * rh) * q;
    const double* t = et + 2 * j;
    const double th = t[0], tl = t[1];
    // 2^(j/64) (1 + rh + b) = th + th rh + (th b + tl (1 + rh + b))
    const double p = th * rh;
    const double pe = fma(th, rh, -p);
    const DD s = fast_two_sum(th, p);
    const double low = ((s.lo + pe) + fma(th, b, tl * (1.0 + (rh + b))));
    return ldexp(s.hi + low, e);
}

// The logarithm pow_pos takes of x: log |x| for finite nonzero x, else 0.
HFV_HD __forceinline__ DD pow_log(double x, const double* lt = pow_log_tab())
{
#ifdef HFV_LIBM_POW
    return DD{0.0, 0.0};  // unused: pow_with_log calls pow()
#endif
    const double ax = fabs(x);
    const bool regular = (ax > 0.0) & (ax < INFINITY);  // false for 0, inf, NaN
    return log_dd(regular ? ax : 1.0, lt);
}

// x^y given l = pow_log(x): the rest of pow_pos, so that powers of one base
// share its logarithm with the same bits as separate pow_pos calls.
HFV_HD __forceinline__ double pow_with_log(double x, double y, DD l, const double* et = pow_exp_tab())
{
#ifdef HFV_LIBM_POW
    return pow(x, y);
#endif
    const double ax = fabs(x);
    const bool regular = (ax > 0.0) & (ax < INFINITY);
    const double wh = y * l.hi;
    const double wl = fma(y, l.hi, -wh) + y * l.lo;
    // beyond these x^y rounds to infinity or to zero
    const bool ovf = wh > 709.782712893384;
    const bool unf = wh < -745.2;
    const bool inrange = !(ovf | unf) & (wh == wh);
    double r = exp_dd(DD{inrange ? wh : 0.0, inrange ? wl : 0.0}, et);
    // a positive finite x other than 1 and a nonzero y with x^y in range:
    // every special case below leaves r as it is
    if (regular & (x > 0.0) & inrange & (x != 1.0) & (y != 0.0)) return r;
    r = ovf ? INFINITY : r;
    r = unf ? 0.0 : r;
    // |x| = 0 or infinity
    r = (ax == 0.0) ? ((y > 0.0) ? 0.0 : INFINITY) : r;
    r = (ax == INFINITY) ? ((y > 0.0) ? INFINITY : 0.0) : r;
    // negative x (-0 and -infinity included): odd integer y keeps the sign,
    // even integer y drops it, any other y gives NaN for finite nonzero x
    const bool yint = rint(y) == y;
    const bool yodd = yint & (rint(0.5 * y) != 0.5 * y);
    const bool neg = signbit(x);
    r = (neg & yodd) ? -r : r;
    r = (neg & !yint & regular) ? NAN : r;
    r = ((x != x) | (y != y)) ? x + y : r;
    // x^0 = 1, 1^y = 1, (-1)^(+-inf) = 1
    r = ((y == 0.0) | (x == 1.0) | ((ax == 1.0) & (fabs(y) == INFINITY))) ? 1.0 : r;
    return r;
}

// x^y for all x, y (C99 pow semantics; accuracy as stated at the top)
HFV_HD __forceinline__ double pow_pos(double x, double y) { return pow_with_log(x, y, pow_log(x)); }

// pow_pos with the log and exp tables read from lt and et (copies of
// POW_LOG_TAB_* and POW_EXP_TAB_*, e.g. in LDS); the same bits
HFV_HD __forceinline__ double pow_pos(double x, double y, const double* lt, const double* et)
{
    return pow_with_log(x, y, pow_log(x, lt), et);
}

}  // namespace hfv
