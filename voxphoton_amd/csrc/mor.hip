// Device primitives of the wavelength sweep by model order reduction
// (Bragg_MOR_script.m, example_Bragg_1500nm_to_1600nm.m:85-114):
//   - weighted sums of operator tensors (lagrange_int / lagrange_int_deriv
//     applied to fN) for vp_op_create_lincomb;
//   - Gram matrices X^H Y and linear combinations X C of device vectors
//     (the projections P' A P, P' b and the lift P x of the reduced model);
//   - per-vector maxima, gathers and scatters of the per-wavelength
//     excitation and field (example_Bragg_1500nm_to_1600nm.m:89-106).
// FP64 complex vectors; sums in a fixed order (no atomics); maxima through
// atomicMax on the bits of non-negative doubles, which is exact.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
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

// out[i] = bits of max_r |X_i[r]|^2 (non-negative doubles order as their bits)
__global__ void __launch_bounds__(KT) k_absmax2(const double2* const* __restrict__ X, int64_t n,
                                               unsigned long long* __restrict__ out) {
  const double2* x = X[blockIdx.y];
  double m = 0.0;
  for (int64_t r = blockIdx.x * int64_t(KT) + threadIdx.x; r < n; r += int64_t(gridDim.x) * KT) {
    const double2 v = x[r];
    m = fmax(m, v.x * v.x + v.y * v.y);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  __shared__ double sm[KT / 64];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < KT / 64; ++w) m = fmax(m, sm[w]);
    atomicMax(out + blockIdx.y, static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Y[i] = alpha w[i] X[idx[i]] (w may be NULL); an index outside [0, nx) gives
// 0 and raises *bad
__global__ void __launch_bounds__(256) k_gather(const double2* __restrict__ X, int64_t nx,
                                               const int64_t* __restrict__ idx, const double2* __restrict__ w,
                                               double2 alpha, int64_t n, double2* __restrict__ Y,
                                               unsigned int* __restrict__ bad) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = idx[i];
    double2 v = make_double2(0.0, 0.0);
    if (r >= 0 && r < nx) {
      const double2 x = X[r];
      double2 a = alpha;
      if (w) {
        const double2 c = w[i];
        a = make_double2(alpha.x * c.x - alpha.y * c.y, alpha.x * c.y + alpha.y * c.x);
      }
      v = make_double2(a.x * x.x - a.y * x.y, a.x * x.y + a.y * x.x);
    } else {
      atomicOr(bad, 1u);
    }
    Y[i] = v;
  }
}

__global__ void __launch_bounds__(256) k_zero_c128(double2* __restrict__ Y, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    Y[i] = make_double2(0.0, 0.0);
}

// Y[idx[i]] = X[i] over a zeroed Y; an index outside [0, ny) raises *bad
__global__ void __launch_bounds__(256) k_scatter(const double2* __restrict__ X, const int64_t* __restrict__ idx,
                                                int64_t n, double2* __restrict__ Y, int64_t ny,
                                                unsigned int* __restrict__ bad) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = idx[i];
    if (r >= 0 && r < ny) Y[r] = X[i];
    else atomicOr(bad, 1u);
  }
}

__global__ void k_fill_u32(unsigned int* __restrict__ p) { *p = 0u; }

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// carves consecutive 256-byte aligned pieces out of the context scratch
struct Carve {
  char* base;
  size_t off = 0;
  template <typename X>
  X* take(size_t count) {
    X* p = reinterpret_cast<X*>(base + off);
    off += align256(count * sizeof(X));
    return p;
  }
};

unsigned grid_for(int64_t n, unsigned cap) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap))); }

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

// Scratch of the allocation-free entry point (vp_vec_dots_noalloc): a
// file-scope device array of 16 MiB that every GPU the library runs on holds
// (its own copy, at its own address) until the process ends.  Host code never
// passes the array's name to a kernel or a copy: it uses the per-device
// address from hipGetSymbolAddress, taken once per device with that device
// current (noalloc_scratch below).
static __device__ __align__(256) char scratch[size_t(16) << 20];

namespace vp {
namespace {

void check_dots_args(vp_ctx ctx, const void* const* X, int k, const void* const* Y, int m, int64_t n,
                     const double* out) {
  VP_REQUIRE(ctx && X && Y && out, ERR_INVALID, "ctx/X/Y/out is NULL");
  VP_REQUIRE(k > 0 && m > 0 && k <= 4096 && m <= 4096 && n > 0, ERR_INVALID, "bad vector counts");
  for (int i = 0; i < k; ++i) VP_REQUIRE(X[i], ERR_INVALID, "X has a NULL vector");
  for (int j = 0; j < m; ++j) VP_REQUIRE(Y[j], ERR_INVALID, "Y has a NULL vector");
}

int dots_blocks(int64_t n) { return int((n + int64_t(KT) * RPT - 1) / (int64_t(KT) * RPT)); }

// scratch of one X^H Y: the two pointer tables, the per-block partial sums
// (every entry written before it is read) and the k x m results
size_t dots_scratch_bytes(int k, int m, int64_t n) {
  return align256(sizeof(void*) * k) + align256(sizeof(void*) * m) +
         align256(size_t(k) * m * dots_blocks(n) * sizeof(double2)) + align256(size_t(k) * m * sizeof(double2));
}

// X^H Y with its scratch at base; returns after the stream drains
void run_dots(hipStream_t s, char* base, const void* const* X, int k, const void* const* Y, int m, int64_t n,
              double* out) {
  const int nblk = dots_blocks(n);
  Carve cv{base};
  auto* dx = cv.take<const double2*>(k);
  auto* dy = cv.take<const double2*>(m);
  auto* part = cv.take<double2>(size_t(k) * m * nblk);
  auto* res = cv.take<double2>(size_t(k) * m);
  VP_HIP(hipMemcpyAsync(dx, X, sizeof(void*) * k, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dy, Y, sizeof(void*) * m, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_vdots, dim3(unsigned(nblk), unsigned(k), unsigned((m + JT - 1) / JT)), dim3(KT), 0, s, dx, dy,
                     k, m, n, part);
  hipLaunchKernelGGL(k_vdots_reduce, dim3(unsigned(k * m)), dim3(KT), 0, s, part, nblk, res);
  VP_HIP(hipGetLastError());
  VP_HIP(hipMemcpyAsync(out, res, size_t(k) * m * sizeof(double2), hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
}

constexpr size_t kNoAllocScratchBytes = sizeof(scratch);
constexpr int kMaxNoAllocDevices = 64;
std::mutex g_noalloc_init_mu;
char* g_noalloc_ptr[kMaxNoAllocDevices] = {};
// calls on one device run one at a time: they share its array
std::mutex g_noalloc_call_mu[kMaxNoAllocDevices];

char* noalloc_scratch(int device) {   // device is current
  std::lock_guard<std::mutex> lk(g_noalloc_init_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    VP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    VP_REQUIRE(p, ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
    g_noalloc_ptr[device] = static_cast<char*>(p);
  }
  return g_noalloc_ptr[device];
}

}  // namespace
}  // namespace vp

using namespace vp;

extern "C" {

int vp_vec_dots(vp_ctx ctx, const void* const* X, int k, const void* const* Y, int m, int64_t n, double* out) {
  VP_API_BEGIN
  check_dots_args(ctx, X, k, Y, m, n, out);
  VP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::mutex> lk(ctx->mu);
  run_dots(ctx->stream, static_cast<char*>(ctx_scratch(ctx, dots_scratch_bytes(k, m, n))), X, k, Y, m, n, out);
  VP_API_END
}

int vp_vec_dots_noalloc(vp_ctx ctx, const void* const* X, int k, const void* const* Y, int m, int64_t n,
                        double* out) {
  VP_API_BEGIN
  check_dots_args(ctx, X, k, Y, m, n, out);
  const size_t need = dots_scratch_bytes(k, m, n);
  VP_REQUIRE(need <= kNoAllocScratchBytes, ERR_UNSUPPORTED,
             "vp_vec_dots_noalloc: k*m*ceil(n/1024) partial sums (" + std::to_string(need) +
                 " bytes with the pointer tables) exceed its " + std::to_string(kNoAllocScratchBytes) +
                 "-byte scratch; use vp_vec_dots");
  VP_REQUIRE(ctx->device >= 0 && ctx->device < kMaxNoAllocDevices, ERR_UNSUPPORTED, "device index too large");
  VP_HIP(hipSetDevice(ctx->device));
  char* base = noalloc_scratch(ctx->device);
  std::lock_guard<std::mutex> lk(g_noalloc_call_mu[ctx->device]);
  run_dots(ctx->stream, base, X, k, Y, m, n, out);
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
  std::lock_guard<std::mutex> lk(ctx->mu);
  const size_t need =
      align256(sizeof(void*) * k) + align256(sizeof(void*) * m) + align256(size_t(k) * m * sizeof(double2));
  Carve cv{static_cast<char*>(ctx_scratch(ctx, need))};
  auto* dx = cv.take<const double2*>(k);
  auto* dy = cv.take<double2*>(m);
  auto* dc = cv.take<double2>(size_t(k) * m);
  VP_HIP(hipMemcpyAsync(dx, X, sizeof(void*) * k, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dy, Y, sizeof(void*) * m, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dc, Cm, size_t(k) * m * sizeof(double2), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_vlincomb, dim3(grid_for(n, 4096), unsigned(m)), dim3(256), 0, s, dx, dc, k, m, n, dy);
  VP_HIP(hipGetLastError());
  VP_HIP(hipStreamSynchronize(s));
  VP_API_END
}

int vp_vec_absmax(vp_ctx ctx, const void* const* X, int k, int64_t n, double* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && X && out, ERR_INVALID, "ctx/X/out is NULL");
  VP_REQUIRE(k > 0 && k <= 4096 && n > 0, ERR_INVALID, "bad vector counts");
  for (int i = 0; i < k; ++i) VP_REQUIRE(X[i], ERR_INVALID, "X has a NULL vector");
  VP_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  std::lock_guard<std::mutex> lk(ctx->mu);
  Carve cv{static_cast<char*>(ctx_scratch(ctx, align256(sizeof(void*) * k) + align256(8 * size_t(k))))};
  auto* dx = cv.take<const double2*>(k);
  auto* res = cv.take<unsigned long long>(k);
  VP_HIP(hipMemcpyAsync(dx, X, sizeof(void*) * k, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill_u64, dim3(unsigned((k + 255) / 256)), dim3(256), 0, s, res, k);
  hipLaunchKernelGGL(k_absmax2, dim3(grid_for(n, 2048), unsigned(k)), dim3(KT), 0, s, dx, n, res);
  VP_HIP(hipGetLastError());
  std::vector<unsigned long long> bits(k);
  VP_HIP(hipMemcpyAsync(bits.data(), res, 8 * size_t(k), hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < k; ++i) {
    double m2;
    std::memcpy(&m2, &bits[i], 8);
    out[i] = std::sqrt(m2);
  }
  VP_API_END
}

int vp_vec_gather(vp_ctx ctx, const void* X, int64_t nx, const int64_t* idx, const void* w, double alpha_re,
                  double alpha_im, int64_t n, void* Y) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && X && idx && Y, ERR_INVALID, "ctx/X/idx/Y is NULL");
  VP_REQUIRE(n > 0 && nx > 0, ERR_INVALID, "bad vector lengths");
  VP_REQUIRE(Y != X, ERR_INVALID, "Y must not alias X");
  VP_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto* bad = static_cast<unsigned int*>(ctx_scratch(ctx, 256));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, s, bad);
  hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 8192)), dim3(256), 0, s, static_cast<const double2*>(X), nx, idx,
                     static_cast<const double2*>(w), make_double2(alpha_re, alpha_im), n, static_cast<double2*>(Y),
                     bad);
  VP_HIP(hipGetLastError());
  unsigned int hbad = 0;
  VP_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  VP_REQUIRE(!hbad, ERR_INVALID, "gather index out of range");
  VP_API_END
}

int vp_vec_scatter(vp_ctx ctx, const void* X, const int64_t* idx, int64_t n, void* Y, int64_t ny) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && X && idx && Y, ERR_INVALID, "ctx/X/idx/Y is NULL");
  VP_REQUIRE(n > 0 && ny > 0, ERR_INVALID, "bad vector lengths");
  VP_REQUIRE(Y != X, ERR_INVALID, "Y must not alias X");
  VP_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto* bad = static_cast<unsigned int*>(ctx_scratch(ctx, 256));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, s, bad);
  hipLaunchKernelGGL(k_zero_c128, dim3(grid_for(ny, 8192)), dim3(256), 0, s, static_cast<double2*>(Y), ny);
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(n, 8192)), dim3(256), 0, s, static_cast<const double2*>(X), idx, n,
                     static_cast<double2*>(Y), ny, bad);
  VP_HIP(hipGetLastError());
  unsigned int hbad = 0;
  VP_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  VP_REQUIRE(!hbad, ERR_INVALID, "scatter index out of range");
  VP_API_END
}

}  // extern "C"
