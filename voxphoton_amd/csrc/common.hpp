// Shared device/host helpers for the VoxPhoton FFT-VIE kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace vp {

template <typename T> struct cplx_t;
template <> struct cplx_t<float> { using type = float2; };
template <> struct cplx_t<double> { using type = double2; };
template <typename T> using C = typename cplx_t<T>::type;

template <typename T>
__host__ __device__ __forceinline__ C<T> mk(T re, T im) { C<T> r; r.x = re; r.y = im; return r; }

__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cneg(float2 a) { return make_float2(-a.x, -a.y); }
__host__ __device__ __forceinline__ double2 cneg(double2 a) { return make_double2(-a.x, -a.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__host__ __device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by -i (fwd=true) or +i
template <bool INV, typename V>
__host__ __device__ __forceinline__ V mul_mi(V a) {
  V r;
  if (INV) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
  return r;
}

template <typename T> __host__ __device__ __forceinline__ C<T> czero() { C<T> r; r.x = T(0); r.y = T(0); return r; }

// Maps the hardware's round-robin placement of workgroups over the 8 XCDs
// to a logical block order in which consecutive logical blocks run on the
// same XCD (and share its L2).  Valid for any total; the tail beyond the
// largest multiple of 8 maps to itself.
__device__ __forceinline__ int xcd_remap(int pid, int total) {
  const int nx = 8;
  int full = (total / nx) * nx;
  if (pid >= full) return pid;
  int per = full / nx;
  return (pid % nx) * per + pid / nx;
}

}  // namespace vp
