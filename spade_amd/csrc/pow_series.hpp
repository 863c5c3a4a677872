// x^n for x near 1 by the binomial series, for the power-law viscosity
// mu_ref (T/T_ref)^n evaluated on every face.
//
// With t = x - 1 (exact for x in [0.5, 2], Sterbenz),
//   x^n = 1 + t P(t) + tail,   P(t) = sum_{k=1..K} C(n, k) t^(k-1),   |t| <= t_max.
// P is evaluated by Estrin's scheme (pairs, then t^2, t^4, t^8 levels: five
// dependent operations instead of thirteen, and no branch, so the chain
// interleaves with the surrounding face arithmetic) and the result is
// fma(P, t, 1).  The coefficients are formed on the host in extended
// precision and rounded once; t_max is the largest power of two <= 1/16 for
// which the tail bound sum_{k>K} |C(n,k)| t_max^k stays below 2^-56.  The
// error is the final rounding plus |t P| <= 1/16 times the few-ulp error of
// P: within one ulp of x^n.  Outside the range the caller uses the library pow.
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
        static_assert(K == 13, "Estrin levels below are written for K = 13");
        const double t = x - 1.0;
        const double t2 = t*t, t4 = t2*t2, t8 = t4*t4;
        // P(t) = sum_{j=0..12} c[j+1] t^j
        const double q0 = __builtin_fma(c[2], t, c[1]);
        const double q1 = __builtin_fma(c[4], t, c[3]);
        const double q2 = __builtin_fma(c[6], t, c[5]);
        const double q3 = __builtin_fma(c[8], t, c[7]);
        const double q4 = __builtin_fma(c[10], t, c[9]);
        const double q5 = __builtin_fma(c[12], t, c[11]);
        const double r0 = __builtin_fma(q1, t2, q0);
        const double r1 = __builtin_fma(q3, t2, q2);
        const double r2 = __builtin_fma(q5, t2, q4);
        const double s0 = __builtin_fma(r1, t4, r0);
        const double s1 = __builtin_fma(c[13], t4, r2);
        const double P  = __builtin_fma(s1, t8, s0);
        return __builtin_fma(P, t, 1.0);
    }
};

} // namespace spd
