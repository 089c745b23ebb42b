// gpis_exp.hpp — exp(x) in double for splat arguments x in [-5, 0] (the 1D noise
// splats exp(-ab^2/2) with |ab| < 3).  Cody-Waite reduction x = k ln2 + r,
// |r| <= ln2/2, then a degree-13 Taylor polynomial by Horner with FMA: relative
// error within 1 ulp of double, so rounding the result to float gives the float
// the reference's (float)exp((double)x) gives (checked against the C library on
// 2e8 arguments by tests/test_exp_splat.py).  About 20 FP64 instructions, without
// the special-case handling of the general-purpose exp.
#pragma once

#include <math.h>

__host__ __device__ __forceinline__ double exp_splat_d(double x) {
    const double k = rint(x * 1.4426950408889634);
    double r = fma(k, -6.93147180369123816490e-01, x);  // ln 2, high part (exact k * hi)
    r = fma(k, -1.90821492927058770002e-10, r);          // ln 2, low part
    double p = 1.0 / 6227020800.0;                        // 1/13!
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
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}
