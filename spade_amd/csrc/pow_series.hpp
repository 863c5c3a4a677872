// x^n for x near 1 by the binomial series, for the power-law viscosity
// mu_ref (T/T_ref)^n evaluated on every face.
//
// With t = x - 1 (exact for x in [0.5, 2], Sterbenz),
//   x^n = sum_{k=0..K} C(n, k) t^k + tail,   |t| <= t_max,
// evaluated by Horner's rule with fused multiply-adds.  The coefficients are
// formed on the host in extended precision and rounded once; t_max is the
// largest power of two <= 1/16 for which the tail bound
// sum_{k>K} |C(n,k)| t_max^k stays below 2^-56.  The result is within one
// ulp of x^n (the final rounding, plus t_max times the few-ulp error of the
// inner polynomial); outside the range the caller uses the library pow.
#pragma once
#include <cmath>

#ifndef SPD_HD
#define SPD_HD __host__ __device__
#endif

namespace spd {

struct PowSeries
{
    static constexpr int K = 13;
    double c[K + 1];     // c[k] = C(n, k), c[0] = 1
    double t_max;        // 0: never in range

    static PowSeries make(double n)
    {
        PowSeries s{};
        long double ck = 1.0L;
        long double coef[K + 41];
        coef[0] = 1.0L;
        for (int k = 1; k < K + 41; ++k)
        {
            ck *= ((long double)n - (long double)(k - 1))/(long double)k;
            coef[k] = ck;
        }
        for (int k = 0; k <= K; ++k) s.c[k] = (double)coef[k];
        s.t_max = 0.0;
        for (double tm = 0.0625; tm >= 1.0/1024.0; tm *= 0.5)
        {
            long double tail = 0.0L, p = 1.0L;
            for (int k = 1; k <= K; ++k) p *= tm;
            for (int k = K + 1; k < K + 41; ++k) { p *= tm; tail += fabsl(coef[k])*p; }
            if (tail <= 1.0L/72057594037927936.0L) { s.t_max = tm; break; }   // 2^-56
        }
        return s;
    }

    SPD_HD bool in_range(double x) const
    {
        const double t = x - 1.0;
        return (t < 0 ? -t : t) <= t_max;
    }

    SPD_HD double eval(double x) const
    {
        const double t = x - 1.0;
        double p = c[K];
        #pragma unroll
        for (int k = K - 1; k >= 1; --k) p = __builtin_fma(p, t, c[k]);
        return __builtin_fma(p, t, 1.0);
    }
};

} // namespace spd
