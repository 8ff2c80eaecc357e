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
// f0*a + f1*b + f2*c as one chain of FMAs per part (5 FMAs + 1 product)
template <typename V>
__host__ __device__ __forceinline__ V cdot3(V f0, V a, V f1, V b, V f2, V c) {
  V r;
  r.x = fma(f0.x, a.x, fma(-f0.y, a.y, fma(f1.x, b.x, fma(-f1.y, b.y, fma(f2.x, c.x, -f2.y * c.y)))));
  r.y = fma(f0.x, a.y, fma(f0.y, a.x, fma(f1.x, b.y, fma(f1.y, b.x, fma(f2.x, c.y, f2.y * c.x)))));
  return r;
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

// Buffer-resource access to a device range of at most 4 GiB: 32-bit element
// offsets (no 64-bit address arithmetic per lane), loads counted on vmcnt
// only (LDS waits never wait on them), and an offset past the range reads 0.
#if defined(__HIP_DEVICE_COMPILE__)
using BufRes = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ BufRes buf_res(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, int(bytes), 0x00020000);
}
// (the raw intrinsics' vector results are bit-cast whole: element access on
// them narrows the load to its first dword)
typedef float vp_f2v __attribute__((ext_vector_type(2)));
typedef double vp_d2v __attribute__((ext_vector_type(2)));
template <typename T> __device__ __forceinline__ C<T> buf_load(BufRes r, uint32_t elem);
template <> __device__ __forceinline__ float2 buf_load<float>(BufRes r, uint32_t elem) {
  const vp_f2v v = __builtin_bit_cast(vp_f2v, __builtin_amdgcn_raw_buffer_load_b64(r, elem * 8u, 0, 0));
  float2 o;
  o.x = v.x;
  o.y = v.y;
  return o;
}
template <> __device__ __forceinline__ double2 buf_load<double>(BufRes r, uint32_t elem) {
  const vp_d2v v = __builtin_bit_cast(vp_d2v, __builtin_amdgcn_raw_buffer_load_b128(r, elem * 16u, 0, 0));
  double2 o;
  o.x = v.x;
  o.y = v.y;
  return o;
}
__device__ __forceinline__ void buf_store(BufRes r, uint32_t elem, float2 a) {
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  vp_f2v v;
  v.x = a.x;
  v.y = a.y;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, elem * 8u, 0, 0);
}
__device__ __forceinline__ void buf_store(BufRes r, uint32_t elem, double2 a) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  vp_d2v v;
  v.x = a.x;
  v.y = a.y;
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, elem * 16u, 0, 0);
}
#else
// host pass: declarations only (device code is never run on the host)
struct BufRes {};
__device__ BufRes buf_res(const void* base, uint32_t bytes);
template <typename T> __device__ C<T> buf_load(BufRes r, uint32_t elem);
__device__ void buf_store(BufRes r, uint32_t elem, float2 a);
__device__ void buf_store(BufRes r, uint32_t elem, double2 a);
#endif

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
