// isph_dual.h -- forward-mode dual numbers for the adjoint kernels.
//
// Dual<N> carries a value and its derivatives along N seed directions.  The
// adjoint kernels evaluate a particle's (or a particle pair's) contribution
// once with the inputs of one side seeded, which gives the local Jacobian of
// that contribution exactly (to rounding); the reverse sweep then contracts
// it with the adjoint of the contribution.  Comparisons and branches act on
// the value, so a piecewise formula is differentiated on the branch the
// value takes.
#pragma once
#include <hip/hip_runtime.h>

namespace isph {

// the double overloads stay visible next to the Dual ones below
using ::fabs;
using ::sqrt;

template <int N>
struct Dual {
    double v;
    double d[N];
    __device__ Dual() {}
    __device__ Dual(double x) : v(x) {
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
};

__device__ __forceinline__ double val(double x) { return x; }
template <int N>
__device__ __forceinline__ double val(const Dual<N>& x) { return x.v; }

// x seeded along direction k
template <int N>
__device__ __forceinline__ Dual<N> seed(double x, int k) {
    Dual<N> r(x);
    r.d[k] = 1.0;
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) {
    Dual<N> r = a;
    r.v = a.v + b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) {
    Dual<N> r = b;
    r.v = a + b.v;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) {
    Dual<N> r = a;
    r.v = a.v - b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) {
    return b * a;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) {
    Dual<N> r;
    r.v = a.v / b;
    const double ib = 1.0 / b;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * ib;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a / b.v;
    const double s = -r.v / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = s * b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) {
    a = a + b;
    return a;
}
template <int N>
__device__ __forceinline__ Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) {
    a = a - b;
    return a;
}

// sqrt with a zero derivative at 0 (the one-sided limit is unbounded; the
// forward formulas only take sqrt(0) where the quantity is identically 0)
template <int N>
__device__ __forceinline__ Dual<N> sqrt(const Dual<N>& a) {
    Dual<N> r;
    r.v = ::sqrt(a.v);
    const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> fabs(const Dual<N>& a) {
    return a.v < 0.0 ? -a : a;
}

#define ISPH_DUAL_CMP(op)                                                                        \
    template <int N>                                                                             \
    __device__ __forceinline__ bool operator op(const Dual<N>& a, const Dual<N>& b) {            \
        return a.v op b.v;                                                                       \
    }                                                                                            \
    template <int N>                                                                             \
    __device__ __forceinline__ bool operator op(const Dual<N>& a, double b) { return a.v op b; } \
    template <int N>                                                                             \
    __device__ __forceinline__ bool operator op(double a, const Dual<N>& b) { return a op b.v; }
ISPH_DUAL_CMP(<)
ISPH_DUAL_CMP(<=)
ISPH_DUAL_CMP(>)
ISPH_DUAL_CMP(>=)
#undef ISPH_DUAL_CMP

}  // namespace isph
