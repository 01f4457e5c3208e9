// Host check of chimera::exp_libm (chimera_amd/csrc/libm_exp.hpp) against the C
// library's exp, built with hipcc as a host program.  Prints one line per argument set:
// "<set> <count> mismatches <m>".
#include <cmath>
#include <cstdint>
#include <cstdio>

#include "libm_exp.hpp"

int main() {
    double (*volatile libm_exp)(double) = std::exp;  // the library call, not a builtin
    // every exponent argument of the Kelvin-Helmholtz set-up (src/mhd_init.f90:341-347)
    const int sizes[] = {4, 64, 100, 256, 384, 1000, 2048};
    for (int n : sizes) {
        const double dx = 1.0 / n;
        long long bad = 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double X = ((double)i + 0.5) * dx, Y = ((double)j + 0.5) * dx;
                const double ex = X - 0.5, ey = Y - 0.25;
                const double t = -(ex * ex + ey * ey) / (2.0 * (0.05 * 0.05));
                if (chimera::exp_libm(t) != libm_exp(t)) ++bad;
            }
        std::printf("kh%d %lld mismatches %lld\n", n, (long long)n * n, bad);
    }
    // random arguments over the supported range: uniform in (-512, 512), and scaled
    // down by 2^-k (k < 64) to cover small |x| down to and below 2^-54
    uint64_t s = 0x9e3779b97f4a7c15ull;
    long long bad = 0, count = 0;
    for (long long m = 0; m < 4000000; ++m) {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        double x = (2.0 * (double)(s >> 11) * 0x1.0p-53 - 1.0) * 511.999;
        if (m % 3 == 1) x = std::ldexp(x, -(int)(s % 64));
        ++count;
        if (chimera::exp_libm(x) != libm_exp(x)) ++bad;
    }
    std::printf("random %lld mismatches %lld\n", count, bad);
    return 0;
}
