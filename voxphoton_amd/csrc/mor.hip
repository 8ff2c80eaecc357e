// Device primitives of the wavelength sweep by model order reduction
// (Bragg_MOR_script.m, example_Bragg_1500nm_to_1600nm.m:85-114):
//   - weighted sums of operator tensors (lagrange_int / lagrange_int_deriv
//     applied to fN) for vp_op_create_lincomb;
//   - Gram matrices X^H Y and linear combinations X C of device vectors
//     (the projections P' A P, P' b and the lift P x of the reduced model).
// FP64 complex vectors; fixed-order reductions (no atomics).
#include <vector>
#include "op.hpp"
#include "capi_common.hpp"

namespace vp {
namespace {

constexpr int KT = 256;     // threads per block
constexpr int RPT = 4;      // rows per thread
constexpr int JT = 8;       // columns of Y per block

template <typename T>
__global__ void __launch_bounds__(256) k_tensor_axpy(C<T>* __restrict__ out, const C<T>* __restrict__ in, double2 w,
                                                    int64_t n, int first) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const C<T> v = in[i];
    const double re = w.x * double(v.x) - w.y * double(v.y);
    const double im = w.x * double(v.y) + w.y * double(v.x);
    const C<T> o = first ? czero<T>() : out[i];
    out[i] = mk<T>(T(double(o.x) + re), T(double(o.y) + im));
  }
}

// partial[p][blk] = sum over the block's rows of conj(X_i[r]) Y_j[r] for the
// pairs p = i + k * j of one i (blockIdx.y) and JT consecutive j (blockIdx.z)
__global__ void __launch_bounds__(KT) k_vdots(const double2* const* __restrict__ X, const double2* const* __restrict__ Y,
                                              int k, int m, int64_t n, double2* __restrict__ partial) {
  const int i = blockIdx.y, j0 = blockIdx.z * JT;
  const int nblk = gridDim.x;
  const double2* xi = X[i];
  double2 acc[JT];
#pragma unroll
  for (int jj = 0; jj < JT; ++jj) acc[jj] = make_double2(0.0, 0.0);
#pragma unroll
  for (int t = 0; t < RPT; ++t) {
    const int64_t r0 = (int64_t(blockIdx.x) * RPT + t) * KT + threadIdx.x;
    const int64_t r = min<int64_t>(r0, n - 1);
    const double2 xv = r0 < n ? xi[r] : make_double2(0.0, 0.0);
#pragma unroll
    for (int jj = 0; jj < JT; ++jj) {
      const double2 yv = Y[min(j0 + jj, m - 1)][r];
      acc[jj].x += xv.x * yv.x + xv.y * yv.y;
      acc[jj].y += xv.x * yv.y - xv.y * yv.x;
    }
  }
  __shared__ double2 sm[KT / 64][JT];
#pragma unroll
  for (int jj = 0; jj < JT; ++jj) {
    for (int off = 32; off > 0; off >>= 1) {
      acc[jj].x += __shfl_xor(acc[jj].x, off);
      acc[jj].y += __shfl_xor(acc[jj].y, off);
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int jj = 0; jj < JT; ++jj) sm[wave][jj] = acc[jj];
  __syncthreads();
  if (threadIdx.x < JT && j0 + int(threadIdx.x) < m) {
    double2 s = make_double2(0.0, 0.0);
    for (int w = 0; w < KT / 64; ++w) { s.x += sm[w][threadIdx.x].x; s.y += sm[w][threadIdx.x].y; }
    partial[(int64_t(i) + int64_t(k) * (j0 + threadIdx.x)) * nblk + blockIdx.x] = s;
  }
}

// out[p] = sum over blocks of partial[p][.] (one workgroup per pair, fixed tree)
__global__ void __launch_bounds__(KT) k_vdots_reduce(const double2* __restrict__ partial, int nblk,
                                                     double2* __restrict__ out) {
  const double2* src = partial + int64_t(blockIdx.x) * nblk;
  double2 s = make_double2(0.0, 0.0);
  for (int b = threadIdx.x; b < nblk; b += KT) { s.x += src[b].x; s.y += src[b].y; }
  for (int off = 32; off > 0; off >>= 1) {
    s.x += __shfl_xor(s.x, off);
    s.y += __shfl_xor(s.y, off);
  }
  __shared__ double2 sm[KT / 64];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t = make_double2(0.0, 0.0);
    for (int w = 0; w < KT / 64; ++w) { t.x += sm[w].x; t.y += sm[w].y; }
    out[blockIdx.x] = t;
  }
}

// Y_j = sum_i X_i C[i + k j] (C on the device, k x m column-major)
__global__ void __launch_bounds__(256) k_vlincomb(const double2* const* __restrict__ X, const double2* __restrict__ Cm,
                                                  int k, int m, int64_t n, double2* const* __restrict__ Y) {
  const int j = blockIdx.y;
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r < n; r += int64_t(gridDim.x) * blockDim.x) {
    double2 s = make_double2(0.0, 0.0);
    for (int i = 0; i < k; ++i) {
      const double2 c = Cm[i + int64_t(k) * j], x = X[i][r];
      s.x += x.x * c.x - x.y * c.y;
      s.y += x.x * c.y + x.y * c.x;
    }
    Y[j][r] = s;
  }
}

}  // namespace

template <typename T>
void tensor_lincomb(C<T>* out, const C<T>* const* in, const double2* w, int k, int64_t n, hipStream_t s) {
  const unsigned blocks = unsigned(std::min<int64_t>((n + 255) / 256, 16384));
  for (int i = 0; i < k; ++i)
    hipLaunchKernelGGL(k_tensor_axpy<T>, dim3(blocks), dim3(256), 0, s, out, in[i], w[i], n, i == 0 ? 1 : 0);
  VP_HIP(hipGetLastError());
}
template void tensor_lincomb<double>(double2*, const double2* const*, const double2*, int, int64_t, hipStream_t);
template void tensor_lincomb<float>(float2*, const float2* const*, const double2*, int, int64_t, hipStream_t);

}  // namespace vp

using namespace vp;

extern "C" {

int vp_vec_dots(vp_ctx ctx, const void* const* X, int k, const void* const* Y, int m, int64_t n, double* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && X && Y && out, ERR_INVALID, "ctx/X/Y/out is NULL");
  VP_REQUIRE(k > 0 && m > 0 && k <= 4096 && m <= 4096 && n > 0, ERR_INVALID, "bad vector counts");
  for (int i = 0; i < k; ++i) VP_REQUIRE(X[i], ERR_INVALID, "X has a NULL vector");
  for (int j = 0; j < m; ++j) VP_REQUIRE(Y[j], ERR_INVALID, "Y has a NULL vector");
  VP_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  DevBuf dx(sizeof(void*) * k), dy(sizeof(void*) * m);
  VP_HIP(hipMemcpyAsync(dx.p, X, dx.bytes, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dy.p, Y, dy.bytes, hipMemcpyHostToDevice, s));
  const int nblk = int((n + int64_t(KT) * RPT - 1) / (int64_t(KT) * RPT));
  DevBuf part(size_t(k) * m * nblk * sizeof(double2)), res(size_t(k) * m * sizeof(double2));
  hipLaunchKernelGGL(k_vdots, dim3(unsigned(nblk), unsigned(k), unsigned((m + JT - 1) / JT)), dim3(KT), 0, s,
                     dx.as<const double2*>(), dy.as<const double2*>(), k, m, n, part.as<double2>());
  hipLaunchKernelGGL(k_vdots_reduce, dim3(unsigned(k * m)), dim3(KT), 0, s, part.as<double2>(), nblk,
                     res.as<double2>());
  VP_HIP(hipGetLastError());
  VP_HIP(hipMemcpyAsync(out, res.p, res.bytes, hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  VP_API_END
}

int vp_vec_lincomb(vp_ctx ctx, const void* const* X, int k, const double* Cm, int m, int64_t n, void* const* Y) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && X && Y && Cm, ERR_INVALID, "ctx/X/C/Y is NULL");
  VP_REQUIRE(k > 0 && m > 0 && k <= 4096 && m <= 4096 && n > 0, ERR_INVALID, "bad vector counts");
  for (int i = 0; i < k; ++i) VP_REQUIRE(X[i], ERR_INVALID, "X has a NULL vector");
  for (int j = 0; j < m; ++j) {
    VP_REQUIRE(Y[j], ERR_INVALID, "Y has a NULL vector");
    for (int i = 0; i < k; ++i) VP_REQUIRE(Y[j] != X[i], ERR_INVALID, "Y must not alias X");
  }
  VP_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  DevBuf dx(sizeof(void*) * k), dy(sizeof(void*) * m), dc(size_t(k) * m * sizeof(double2));
  VP_HIP(hipMemcpyAsync(dx.p, X, dx.bytes, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dy.p, Y, dy.bytes, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dc.p, Cm, dc.bytes, hipMemcpyHostToDevice, s));
  const unsigned blocks = unsigned(std::min<int64_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(k_vlincomb, dim3(blocks, unsigned(m)), dim3(256), 0, s, dx.as<const double2*>(),
                     dc.as<double2>(), k, m, n, dy.as<double2*>());
  VP_HIP(hipGetLastError());
  VP_HIP(hipStreamSynchronize(s));
  VP_API_END
}

}  // extern "C"
