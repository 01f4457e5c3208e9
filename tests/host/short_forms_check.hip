// Bit-for-bit check of the stage kernel's short FP64 division and square root
// (recip_nr / div_nr / sqrt_nr / recip_uniform in mhd_phys.hpp) against hipcc's
// correctly rounded a / b and sqrt(x), over random operands inside the ranges where the
// short forms are used.  Built into tests/host/libshort_forms_check.so by the package
// build; tests/test_gpu_short_forms.py drives it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mhd_phys.hpp"

using namespace chimera;

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// A double with binary exponent in [emin, emax] and a mantissa drawn from r; every
// fourth draw takes an edge mantissa (all zeros, all ones, or a single low bit).
__device__ __forceinline__ double draw(uint64_t r, int emin, int emax, bool sign) {
    const int span = emax - emin + 1;
    const int e = emin + (int)((r >> 52) % (uint64_t)span);
    uint64_t m = r & 0xfffffffffffffull;
    switch ((r >> 61) & 7) {
        case 0: m = 0; break;
        case 1: m = 0xfffffffffffffull; break;
        case 2: m = 1; break;
        default: break;
    }
    const uint64_t bits = ((uint64_t)(e + 1023) << 52) | m | (sign ? (1ull << 63) : 0ull);
    return __longlong_as_double((long long)bits);
}

__device__ __forceinline__ bool same_bits(double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b);
}

// mode 0: quotients a / b, a in [2^ea0, 2^ea1], b in [2^eb0, 2^eb1] (random signs when
//         sign != 0), kept where |a / b| lies in [2^-900, 2^700];
// mode 1: sqrt(x), x in [2^ea0, 2^ea1];
// mode 2: a / b with a wave-uniform b (recip_uniform), as for P_th / (g - 1);
// mode 3: exact quotients a = b k (k a small integer), where a correction step that
//         goes wrong shows at once.
__global__ void check_kernel(uint64_t seed, long long n, int mode, int ea0, int ea1, int eb0,
                             int eb1, int sign, unsigned long long *bad,
                             unsigned long long *used) {
    unsigned long long nbad = 0, nused = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t r1 = splitmix64(seed ^ (uint64_t)(2 * i));
        const uint64_t r2 = splitmix64(seed ^ (uint64_t)(2 * i + 1));
        if (mode == 1) {
            const double x = draw(r1, ea0, ea1, false);
            ++nused;
            if (!same_bits(sqrt_nr(x), sqrt(x))) ++nbad;
            continue;
        }
        double a, b;
        if (mode == 2) {
            // one divisor per wave: the wave's first lane's draw
            const uint64_t rb = splitmix64(seed ^ 0x5bd1e995ull ^ (uint64_t)(i / 64));
            b = draw(rb, eb0, eb1, false);
            a = draw(r1, ea0, ea1, sign && (r2 & 1));
        } else if (mode == 3) {
            b = draw(r2, eb0, eb1, sign && (r2 & 1));
            a = b * (double)(1 + (int)(r1 % 1000));
        } else {
            a = draw(r1, ea0, ea1, sign && (r2 & 1));
            b = draw(r2, eb0, eb1, sign && (r2 & 2));
        }
        const double q = a / b;
        const double aq = fabs(q);
        if (!(aq >= 0x1p-900 && aq <= 0x1p700)) continue;
        ++nused;
        const double s = mode == 2 ? div_nr(a, recip_uniform(b)) : div_nr(a, recip_nr(b));
        if (!same_bits(s, q)) ++nbad;
    }
    atomicAdd(bad, nbad);
    atomicAdd(used, nused);
}

}  // namespace

// Returns 0 on success (1000 + the HIP error code when the counters cannot be
// allocated); *bad counts operands whose short form differs from the IEEE
// result, *used the operands checked.  Synchronous.
extern "C" int short_forms_check(unsigned long long seed, long long n, int mode, int ea0,
                                 int ea1, int eb0, int eb1, int sign,
                                 unsigned long long *bad, unsigned long long *used) {
    unsigned long long *d = nullptr;
    hipError_t e = hipMalloc(&d, 2 * sizeof(unsigned long long));
    if (e != hipSuccess) return 1000 + (int)e;
    int rc = 0;
    if (hipMemset(d, 0, 2 * sizeof(unsigned long long)) != hipSuccess) rc = 2;
    if (!rc) {
        // mode 2 needs whole waves with one divisor each: 256-thread blocks, n a
        // multiple of 64 (the caller's sizes are)
        hipLaunchKernelGGL(check_kernel, dim3(1024), dim3(256), 0, 0, (uint64_t)seed, n, mode,
                           ea0, ea1, eb0, eb1, sign, d, d + 1);
        if (hipGetLastError() != hipSuccess) rc = 3;
    }
    unsigned long long h[2] = {0, 0};
    if (!rc && hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) rc = 4;
    hipFree(d);
    *bad = h[0];
    *used = h[1];
    return rc;
}
