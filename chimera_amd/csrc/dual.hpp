// Forward-mode dual numbers for the Riemann-solver Jacobians of the adjoint.
//
// Dual<N> carries a value and N directional derivatives.  Comparisons act on the
// value only, so a solver instantiated with Dual takes exactly the branches the FP64
// forward pass took (same operations in the same order give the same values) and
// returns the derivative of the branch taken -- the adjoint differentiates through
// the limiter / MOOD / cap / catch branches as they were taken.
#pragma once

#include <hip/hip_runtime.h>

namespace chimera {

template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
__device__ __forceinline__ Dual<N> dconst(double v) {
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
}

#define CH_DUAL_BIN(OP, VEXPR, DEXPR)                                                    \
    template <int N>                                                                     \
    __device__ __forceinline__ Dual<N> operator OP(const Dual<N> &a, const Dual<N> &b) { \
        Dual<N> r;                                                                       \
        r.v = VEXPR;                                                                     \
        _Pragma("unroll") for (int k = 0; k < N; ++k) r.d[k] = DEXPR;                    \
        return r;                                                                        \
    }
CH_DUAL_BIN(+, a.v + b.v, a.d[k] + b.d[k])
CH_DUAL_BIN(-, a.v - b.v, a.d[k] - b.d[k])
CH_DUAL_BIN(*, a.v * b.v, a.d[k] * b.v + a.v * b.d[k])
#undef CH_DUAL_BIN

template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v / b.v;
    const double inv = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a) {
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}

// mixed double / Dual arithmetic
template <int N> __device__ __forceinline__ Dual<N> operator+(double a, const Dual<N> &b) { return dconst<N>(a) + b; }
template <int N> __device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, double b) { return a + dconst<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> operator-(double a, const Dual<N> &b) { return dconst<N>(a) - b; }
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, double b) { return a - dconst<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> operator*(double a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
    return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, double b) {
    Dual<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, double b) { return a / dconst<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> operator/(double a, const Dual<N> &b) { return dconst<N>(a) / b; }

template <int N> __device__ __forceinline__ bool operator<(const Dual<N> &a, const Dual<N> &b) { return a.v < b.v; }
template <int N> __device__ __forceinline__ bool operator>(const Dual<N> &a, const Dual<N> &b) { return a.v > b.v; }
template <int N> __device__ __forceinline__ bool operator<=(const Dual<N> &a, const Dual<N> &b) { return a.v <= b.v; }
template <int N> __device__ __forceinline__ bool operator>=(const Dual<N> &a, const Dual<N> &b) { return a.v >= b.v; }
template <int N> __device__ __forceinline__ bool operator<(const Dual<N> &a, double b) { return a.v < b; }
template <int N> __device__ __forceinline__ bool operator>(const Dual<N> &a, double b) { return a.v > b; }
template <int N> __device__ __forceinline__ bool operator<=(const Dual<N> &a, double b) { return a.v <= b; }
template <int N> __device__ __forceinline__ bool operator>=(const Dual<N> &a, double b) { return a.v >= b; }

template <int N>
__device__ __forceinline__ Dual<N> dsqrt(const Dual<N> &a) {
    Dual<N> r;
    r.v = sqrt(a.v);
    const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dabs(const Dual<N> &a) {
    return a.v < 0.0 ? -a : (a.v > 0.0 ? a : dconst<N>(0.0) + a * 0.0);
}
// fmax / fmin keep the argument the FP64 forward pass returns (the first on ties).
template <int N>
__device__ __forceinline__ Dual<N> dmax(const Dual<N> &a, const Dual<N> &b) { return b.v > a.v ? b : a; }
template <int N>
__device__ __forceinline__ Dual<N> dmin(const Dual<N> &a, const Dual<N> &b) { return b.v < a.v ? b : a; }
template <int N>
__device__ __forceinline__ Dual<N> dmax(const Dual<N> &a, double b) { return b > a.v ? dconst<N>(b) : a; }
template <int N>
__device__ __forceinline__ Dual<N> dmin(const Dual<N> &a, double b) { return b < a.v ? dconst<N>(b) : a; }
template <int N>
__device__ __forceinline__ bool dfinite(const Dual<N> &a) { return __builtin_isfinite(a.v); }
template <int N>
__device__ __forceinline__ double dsign1(const Dual<N> &a) { return copysign(1.0, a.v); }

}  // namespace chimera
