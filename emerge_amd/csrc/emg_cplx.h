// Minimal complex helpers over HIP vector types (float2 / double2).
#pragma once

#include <hip/hip_runtime.h>

namespace emg {

template <typename R>
struct CT;
template <>
struct CT<float> {
    using type = float2;
};
template <>
struct CT<double> {
    using type = double2;
};
template <typename R>
using cplx = typename CT<R>::type;

template <typename R>
__host__ __device__ __forceinline__ cplx<R> cmk(R re, R im) {
    cplx<R> z;
    z.x = re;
    z.y = im;
    return z;
}
template <typename T>
__host__ __device__ __forceinline__ T cadd(T a, T b) {
    T z;
    z.x = a.x + b.x;
    z.y = a.y + b.y;
    return z;
}
template <typename T>
__host__ __device__ __forceinline__ T csub(T a, T b) {
    T z;
    z.x = a.x - b.x;
    z.y = a.y - b.y;
    return z;
}
template <typename T>
__host__ __device__ __forceinline__ T cmul(T a, T b) {
    T z;
    z.x = a.x * b.x - a.y * b.y;
    z.y = a.x * b.y + a.y * b.x;
    return z;
}
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ T cmulc(T a, T b) {
    T z;
    z.x = a.x * b.x + a.y * b.y;
    z.y = a.x * b.y - a.y * b.x;
    return z;
}
// acc + a*b
template <typename T>
__host__ __device__ __forceinline__ T cfma(T a, T b, T acc) {
    T z;
    z.x = acc.x + a.x * b.x - a.y * b.y;
    z.y = acc.y + a.x * b.y + a.y * b.x;
    return z;
}
template <typename T>
__host__ __device__ __forceinline__ double cabs2(T a) {
    return double(a.x) * double(a.x) + double(a.y) * double(a.y);
}
__host__ __device__ __forceinline__ double2 to_d(float2 a) { return make_double2(a.x, a.y); }
__host__ __device__ __forceinline__ double2 to_d(double2 a) { return a; }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> from_d(double2 a);
template <>
__host__ __device__ __forceinline__ float2 from_d<float>(double2 a) {
    return make_float2(float(a.x), float(a.y));
}
template <>
__host__ __device__ __forceinline__ double2 from_d<double>(double2 a) {
    return a;
}

}  // namespace emg
