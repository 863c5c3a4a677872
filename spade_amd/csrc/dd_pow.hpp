// x^n for x near 1, evaluated in double-double and rounded once.
//
// Used for the power-law viscosity mu_ref (T/T_ref)^n on every face, where
// T/T_ref stays close to 1.  Method: with t = x - 1 (exact for x in [0.5, 2]),
//   log(x) = t - t^2/2 + t^3 P(t)         (Taylor, |t| <= 1/32, 13 terms)
//   x^n    = exp(y),  y = n log(x)        (y as a double-double)
//   exp(y) = 1 + y + y^2/2 + y^3 Q(y)     (Taylor, |y| <= 1/32, 9 terms)
// The leading terms are carried as double-double (error-free products via
// fma, Fast2Sum for the sums) and the tails in double, so the value before
// the final rounding has a relative error near 2^-63: the result is the
// correctly rounded x^n except in about one case in a few hundred.  Outside
// |t| <= 1/32 or |y| <= 1/32 the caller falls back to the library pow.
#pragma once

#ifndef SPD_HD
#define SPD_HD __host__ __device__
#endif

namespace spd {

struct PowNearOne
{
    // true when x^n can be evaluated by dd_pow_near_one
    SPD_HD static inline bool in_range(double x, double n)
    {
        const double t = x - 1.0;
        const double at = t < 0 ? -t : t;
        const double an = n < 0 ? -n : n;
        return at <= 0.03125 && an*at*1.05 <= 0.03125;
    }
};

SPD_HD inline double dd_pow_near_one(double x, double n)
{
    const double t = x - 1.0;
    // t^2 exactly as hi + lo
    const double t2 = t*t;
    const double t2l = __builtin_fma(t, t, -t2);
    // P(t) = sum_{k>=3} (-1)^(k+1) t^(k-3) / k, k = 3..15
    double P = -1.0/15.0;
    P = __builtin_fma(P, t,  1.0/14.0);
    P = __builtin_fma(P, t, -1.0/13.0);
    P = __builtin_fma(P, t,  1.0/12.0);
    P = __builtin_fma(P, t, -1.0/11.0);
    P = __builtin_fma(P, t,  1.0/10.0);
    P = __builtin_fma(P, t, -1.0/9.0);
    P = __builtin_fma(P, t,  1.0/8.0);
    P = __builtin_fma(P, t, -1.0/7.0);
    P = __builtin_fma(P, t,  1.0/6.0);
    P = __builtin_fma(P, t, -1.0/5.0);
    P = __builtin_fma(P, t,  1.0/4.0);
    P = __builtin_fma(P, t, -1.0/3.0);
    const double R = -(t2*t)*P;                 // t^3/3 - t^4/4 + ...
    // log = t + (-t^2/2) + R, as (lh, ll)
    const double h = -0.5*t2, hl = -0.5*t2l;    // exact scalings
    const double lh = t + h;                    // |h| <= |t|: Fast2Sum
    const double ll0 = (h - (lh - t)) + (hl + R);
    const double L = lh + ll0;                  // renormalise: |ll0| << |lh|
    const double ll = ll0 - (L - lh);
    // y = n log(x) as (yh, yl), |yl| <= ulp(yh)
    const double yh = n*L;
    const double yl = __builtin_fma(n, L, -yh) + n*ll;
    // exp(yh + yl) = 1 + e, e = yh + yh^2/2 + yh^3 Q(yh) + yl (1 + yh)
    const double y2 = yh*yh;
    const double y2l = __builtin_fma(yh, yh, -y2);
    double Q = 1.0/362880.0;
    Q = __builtin_fma(Q, yh, 1.0/40320.0);
    Q = __builtin_fma(Q, yh, 1.0/5040.0);
    Q = __builtin_fma(Q, yh, 1.0/720.0);
    Q = __builtin_fma(Q, yh, 1.0/120.0);
    Q = __builtin_fma(Q, yh, 1.0/24.0);
    Q = __builtin_fma(Q, yh, 1.0/6.0);
    const double z = 0.5*y2, zl = 0.5*y2l;
    const double eh = yh + z;                   // |z| <= |yh|: Fast2Sum
    const double el = (z - (eh - yh)) + (zl + __builtin_fma((y2*yh), Q, __builtin_fma(yl, yh, yl)));
    // 1 + eh + el
    const double s = 1.0 + eh;                  // |eh| < 1: Fast2Sum
    const double sl = (eh - (s - 1.0)) + el;
    return s + sl;
}

} // namespace spd
