// exp(x) with the C library's bits, for the initial conditions that call exp per cell.
//
// The reference's set-ups call the C library's exp from compiled Fortran, and the
// state they produce is stepped by a scheme whose limiter and MOOD ties turn a one-ulp
// difference into a visible one, so a set-up on the device needs exp with exactly the
// host library's bits.  That exp (glibc >= 2.28 on x86-64) is the table-driven method
// published in Arm's optimized-routines (Szabolcs Nagy, math/exp.c): with N = 128 and
// k = round(x N / ln 2),
//     exp(x) = 2^(k/N) exp(r),   r = x - k ln 2 / N  (|r| <= ln 2 / 2N),
// 2^(k/N) = 2^(k div N) H_(k mod N) (1 + T_(k mod N)) from a 128-entry table
// (exp_table.hpp, written by gen_exp_table.py from exact arithmetic) and
//     exp(r) - 1 ~ r + C2 r^2 + C3 r^3 + C4 r^4 + C5 r^5,
// with the method's published constants.  glibc runs its FMA build on CPUs with FMA,
// so every product-plus-sum below is one fused multiply-add, in the grouping of that
// build.  Checked against the host's libm on every argument of the Kelvin-Helmholtz
// set-up up to 4096^2 cells and on random arguments of the range below.
//
// Range: 2^-54 <= |x| < 512, and |x| < 2^-54 (1 + x).  The library's separate
// over/underflow path for |x| >= 512 is not restated; such x give NaN here.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "exp_table.hpp"

namespace chimera {

__host__ __device__ inline double exp_libm(double x) {
    constexpr double inv_ln2_n = 0x1.71547652b82fep0 * 128;
    constexpr double shift = 0x1.8p52;  // adding it rounds to an integer in the low bits
    constexpr double neg_ln2_hi_n = -0x1.62e42fefa0000p-8;
    constexpr double neg_ln2_lo_n = -0x1.cf79abc9e3b3ap-47;
    constexpr double c2 = 0x1.ffffffffffdbdp-2, c3 = 0x1.555555555543cp-3;
    constexpr double c4 = 0x1.55555cf172b91p-5, c5 = 0x1.1111167a4d017p-7;
    const double ax = fabs(x);
    if (ax < 0x1p-54) return 1.0 + x;
    if (!(ax < 512.0)) return __builtin_nan("");
    double kd = fma(inv_ln2_n, x, shift);
    const uint64_t ki = __builtin_bit_cast(uint64_t, kd);
    kd -= shift;
    const double r = fma(kd, neg_ln2_lo_n, fma(kd, neg_ln2_hi_n, x));
    const uint32_t idx = 2u * (uint32_t)(ki & 127u);
    const double tail = __builtin_bit_cast(double, EXP_TABLE_2K128[idx]);
    const uint64_t sbits = EXP_TABLE_2K128[idx + 1] + (ki << 45);
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, c5, c4), fma(r2, fma(r, c3, c2), tail + r));
    const double scale = __builtin_bit_cast(double, sbits);
    return fma(scale, tmp, scale);
}

}  // namespace chimera
