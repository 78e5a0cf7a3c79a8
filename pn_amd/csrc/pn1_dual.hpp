// pn1_dual.hpp -- forward-mode dual numbers for the adjoint kernels.
//
// Dual<N> carries a value and N tangents.  The adjoint kernels evaluate a
// scalar (the cotangent-weighted acceleration terms of one pair or one
// star) with the lane's own coordinates seeded as unit tangents, so the
// tangents of the result are that scalar's gradient with respect to them.
// Plain doubles mix freely with duals (a double is a constant).  The
// translation unit that uses this header is compiled with
// -ffinite-math-only -fno-signed-zeros, so products with the zero tangents
// of seeds fold away at compile time.
#pragma once
#include <hip/hip_runtime.h>

#include "pn1_common.hpp"

namespace pn1dev {

template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
__host__ __device__ __forceinline__ Dual<N> dseed(double v, int k)
{
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = i == k ? 1.0 : 0.0;
    return r;
}

__host__ __device__ __forceinline__ double val(double a) { return a; }
template <int N>
__host__ __device__ __forceinline__ double val(const Dual<N>& a) { return a.v; }

// ---- addition / subtraction ----------------------------------------------
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b)
{
    Dual<N> r = a;
    r.v = a.v + b;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b)
{
    return b + a;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a)
{
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b)
{
    Dual<N> r = a;
    r.v = a.v - b;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -b.d[i];
    return r;
}

// ---- multiplication / division -------------------------------------------
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = __builtin_fma(a.v, b.d[i], a.d[i] * b.v);
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b)
{
    Dual<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b)
{
    return b * a;
}
// 1 / b
template <int N>
__host__ __device__ __forceinline__ Dual<N> dinv(const Dual<N>& b)
{
    Dual<N> r;
    r.v = 1.0 / b.v;
    const double c = -r.v * r.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = c * b.d[i];
    return r;
}
__host__ __device__ __forceinline__ double dinv(double b) { return 1.0 / b; }

// b^(-1/2) (b > 0), through the library's FP64 rsqrt
template <int N>
__host__ __device__ __forceinline__ Dual<N> drsq(const Dual<N>& b)
{
    Dual<N> r;
    r.v = rsq64(b.v);
    const double c = -0.5 * r.v * r.v * r.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = c * b.d[i];
    return r;
}
__host__ __device__ __forceinline__ double drsq(double b) { return rsq64(b); }

}  // namespace pn1dev
