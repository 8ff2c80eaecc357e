// Operator / context objects behind the opaque C handles.
#pragma once
#include <atomic>
#include <memory>
#include <mutex>
#include <vector>
#include "voxphoton.h"
#include "vp_internal.hpp"

struct vp_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // grow-only device scratch of the context's vector entry points (pointer
  // tables, partial sums, flags); held under mu for a whole call
  std::mutex mu;
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // recycled Krylov space of mixed-precision solves: at most this many
  // harmonic Ritz directions of the first inner solve's last cycle (0: the
  // whole cycle; -1: VP_IR_RECYCLE_K or the whole cycle)
  std::atomic<int> recycle_width{-1};
};

namespace vp {

// Device memory of the library.  Requests of 64 MiB and more go through a
// per-device cache: a released buffer is kept and handed to a later request
// of between 4/5 of its size and its size once the work enqueued before its
// release on every stream of the library has finished (events, no device
// synchronisation); 48 GiB at most are kept, and the cache is emptied and the
// allocation retried when hipMalloc runs out of memory.  Repeated solves of
// one size then allocate nothing.
void* dev_alloc(size_t bytes);
void dev_free(void* p);
void dev_cache_flush();
// The library's streams: a released cached buffer is reused only after the
// work enqueued on each of them before the release has finished.
void register_stream(int device, hipStream_t s);
void unregister_stream(hipStream_t s);

// RAII device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  void alloc(size_t b) {
    release();
    if (b) p = dev_alloc(b);
    bytes = b;
  }
  void release() {
    if (p) dev_free(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  template <typename X> X* as() const { return reinterpret_cast<X*>(p); }
};

// RAII pinned host buffer (hipHostMalloc).
struct PinnedBuf {
  void* p = nullptr;
  explicit PinnedBuf(size_t b, unsigned flags = hipHostMallocDefault) { VP_HIP(hipHostMalloc(&p, b, flags)); }
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
};

// Per-precision device state of an operator.
template <typename T>
struct OpPrec {
  bool present = false;
  Geometry g{};     // per precision: residue classes depend on the LDS budget
  DevBuf fN;        // class layout, prescaled
  DevBuf mc, mr;    // per unknown voxel (structured) or unknown (generic)
  DevBuf tw;        // twiddles: Lp | Mp | Np entries
  DevBuf work;      // spectrum workspace 3 * Lp * M * N
  DevBuf work2;     // split variant: y-transformed workspace 3 * Lp * Mp * N
  DevBuf xs, ys;    // vector staging for host-memory calls
  DevBuf probs;     // one ProbPtrs<T>
  T gram = T(1);
  MatvecPlan plan{};
  const C<T>* twx() const { return tw.as<C<T>>(); }
  const C<T>* twy() const { return tw.as<C<T>>() + g.x.P; }
  const C<T>* twz() const { return tw.as<C<T>>() + g.x.P + g.y.P; }
};

}  // namespace vp

struct vp_op_s {
  vp_ctx ctx = nullptr;
  int L = 0, M = 0, N = 0;
  // FFT sizes the caller fixed (its fN's, or requested explicitly); 0: each
  // precision's plan takes the autotuner's sizes (OpPrec::g.{x,y,z}.P)
  int Lp = 0, Mp = 0, Np = 0;
  int l0 = 0, m0 = 0, n0 = 0, Lb = 0, Mb = 0, Nb = 0;   // bounding box of the unknowns' voxels
  int64_t V = 0, nS = 0;
  int structured = 0;
  int64_t n = 0;               // unknowns
  double gram = 1.0;           // FP64 identity weight
  double scale32 = 1.0;        // FP32 operator = A / scale32
  bool has_mr = false;
  bool fsym = false;           // tensor is parity symmetric: split variant stores its octant
  bool unit_contrast = false;  // Mc = 1 at every unknown and no Mr: A = Gram I - N (field evaluation)
  bool all_voxels = false;     // idx = 1..3V in order: unknown vector = the whole L x M x N x 3 field
  vp::DevBuf pos;              // voxel map
  vp::OpPrec<double> p64;
  vp::OpPrec<float> p32;
};

namespace vp {

template <typename T> OpPrec<T>& prec_of(vp_op_s* op);
template <> inline OpPrec<double>& prec_of<double>(vp_op_s* op) { return op->p64; }
template <> inline OpPrec<float>& prec_of<float>(vp_op_s* op) { return op->p32; }

// Build helpers (setup.hip)
// Chooses residue classes and launch shapes for one precision at FFT sizes
// Lp x Mp x Np (each at least 2B-1 for the unknowns' box B); throws
// ERR_UNSUPPORTED when no shape fits.
void build_geometry(Geometry& g, MatvecPlan& mp, const vp_op_s& op, int Lp, int Mp, int Np, bool fp64, int device,
                    int variant);
template <typename T> void upload_twiddles(OpPrec<T>& P, hipStream_t s);
// Writes the class-layout tensor of geometry g from a true-frequency device
// tensor F (Lp x Mp x Np x 6 complex128, column-major), times scale.
template <typename T> void layout_fN(const Geometry& g, const double2* F, DevBuf& out, double scale, hipStream_t s);
// Octant layout [kxc][kzc][comp][kyc] of a parity-symmetric tensor (split variant).
template <typename T>
void layout_fN_sym(const Geometry& g, const double2* F, DevBuf& out, double scale, hipStream_t s);
// True when F(P - k) = parity * F(k) along every axis (to 1e-12 of max |F|).
bool fN_parity_symmetric(int Lp, int Mp, int Np, const double2* F, hipStream_t s);
// True-frequency tensor F (Lp x Mp x Np x 6) of the circulant embedding of a
// device Toeplitz tensor G (L x M x N x 6, offsets 0..D-1 per axis); sizes at
// least 2D-1.
void toeplitz_to_fN(int L, int M, int N, int Lp, int Mp, int Np, int device, const double2* G, double2* F,
                    hipStream_t s);
// A caller's true-frequency tensor F (P0[0] x P0[1] x P0[2] x 6) resampled
// to the circulant sizes P1 (each at least 2B-1 for the unknowns' box
// extents B): the same operator on the box.  Returns false (out untouched)
// when no size changes or a changed length does not plan or fit LDS.
bool resample_fN(const double2* F, const int P0[3], const int B[3], const int P1[3], int device, DevBuf& out,
                 hipStream_t s);
// Device assembly of the Toeplitz tensor.
// ko_im < 0 assembles a damped kernel e^{-j(ko + j ko_im)R}/R (preconditioners only).
void assemble_toeplitz_dev(int L, int M, int N, double dx, double ko, double2* G, hipStream_t s, double ko_im = 0.0);

template <typename T>
void op_apply_dev(vp_op_s* op, int transp, const C<T>* x, C<T>* y);

// idx (1-based, length n) -> voxel map; structured = [idxS; V+idxS; 2V+idxS] form
void index_map(int64_t V, const int64_t* idx, int64_t n, std::vector<int>& pos, int& structured, int64_t& nS);

// out = sum_i w[i] in[i] over n entries (tensor layouts of one geometry; mor.hip)
// ctx->scratch grown to at least bytes (caller holds ctx->mu)
void* ctx_scratch(vp_ctx_s* ctx, size_t bytes);

template <typename T>
void tensor_lincomb(C<T>* out, const C<T>* const* in, const double2* w, int k, int64_t n, hipStream_t s);

// Generic device helpers (setup.hip)
void convert_c128_to_c64(const double2* in, float2* out, int64_t n, double scale, hipStream_t s);
void convert_c64_to_c128(const float2* in, double2* out, int64_t n, double scale, hipStream_t s);

}  // namespace vp
