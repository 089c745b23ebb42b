// Host check of exp_splat_d (gpis_extend_amd/csrc/gpis_exp.hpp) against the C library:
// the float roundings must agree on every argument tried.  Built and run by
// tests/test_exp_splat.py.
#define __host__
#define __device__
#define __forceinline__ inline
#include <cmath>
#include <cstdio>
#include <random>
#include "../gpis_extend_amd/csrc/gpis_exp.hpp"

int main(int argc, char **argv) {
    long long n = argc > 1 ? atoll(argv[1]) : 20000000;
    std::mt19937_64 g(7);
    std::uniform_real_distribution<float> U(-4.5f, 0.0f);
    long long bad = 0;
    double maxrel = 0;
    for (long long i = 0; i < n; ++i) {
        // random arguments, and the splat form -(ab*ab/2) of random impulse offsets
        float x = (i & 1) ? U(g) : -([&] { float ab = 3.0f * (U(g) / 4.5f + 0.5f) * 2.0f; return ab * ab / 2.0f; }());
        double a = exp_splat_d((double)x), b = std::exp((double)x);
        double rel = std::fabs(a - b) / b;
        if (rel > maxrel) maxrel = rel;
        if ((float)a != (float)b) ++bad;
    }
    printf("%lld %lld %.3g\n", n, bad, maxrel);
    return 0;
}
