// Host check of two numerical building blocks of the FP64 flux kernel:
//   1. the binomial-series x^n (PowSeries) against x^n in quad precision;
//   2. division by a constant with Markstein's correction step against IEEE '/'.
// Prints "pow_samples pow_misrounded max_err_ulp div_samples div_mismatch
// exponents_with_full_range".
#include <cstdio>
#include <cmath>
#include <random>
#include <quadmath.h>
#define SPD_HD
#include "../../spade_amd/csrc/pow_series.hpp"

int main()
{
    std::mt19937_64 rng(2024);
    long pn = 0, pbad = 0; double maxulp = 0.0;
    const double ns[] = {0.76, 0.7, 0.5, 1.5, -0.5, 2.0/3.0};
    long prange = 0;
    for (double n : ns)
    {
        const spd::PowSeries ps = spd::PowSeries::make(n);
        if (ps.t_max >= 0.0625) ++prange;
        std::uniform_real_distribution<double> U(1.0 - ps.t_max, 1.0 + ps.t_max);
        for (int s = 0; s < 200000; ++s)
        {
            double x = U(rng);
            if (s % 3 == 0) x = 1.0 + (U(rng) - 1.0)*1e-3;
            if (!ps.in_range(x)) continue;
            ++pn;
            const double a = ps.eval(x);
            const __float128 q = powq((__float128)x, (__float128)n);
            if (a != (double)q) ++pbad;
            const double ulp = std::nextafter((double)q, 2.0) - (double)q;
            const double e = (double)fabsq(((__float128)a - q)/(__float128)ulp);
            if (e > maxulp) maxulp = e;
        }
    }
    long dn = 0, dbad = 0;
    const double bs[] = {0.72, 0.7, 0.9, 1.1, 0.123456789, 7.0/9.0};
    std::uniform_real_distribution<double> E(-40, 10);
    for (double b : bs)
    {
        const double y = 1.0/b;
        for (int s = 0; s < 250000; ++s)
        {
            double a = std::exp2(E(rng))*(1 + std::ldexp((double)(rng() >> 11), -53));
            if (s & 1) a = -a;
            const double q0 = a*y;
            const double r = std::fma(-q0, b, a);
            const double q = std::fma(r, y, q0);
            ++dn;
            if (q != a/b) ++dbad;
        }
    }
    std::printf("%ld %ld %.6f %ld %ld %ld\n", pn, pbad, maxulp, dn, dbad, prange);
    return 0;
}
