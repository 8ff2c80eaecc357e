// C ABI: library, contexts, device memory helpers and the N operator.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>
#include "voxphoton.h"
#include "capi_common.hpp"
#include "op.hpp"

namespace vp {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace vp

using namespace vp;


namespace vp {
void* ctx_scratch(vp_ctx_s* ctx, size_t bytes) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) VP_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->scratch) dev_free(ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    const size_t b = std::max<size_t>(bytes, size_t(1) << 20);
    ctx->scratch = dev_alloc(b);
    ctx->scratch_bytes = b;
  }
  return ctx->scratch;
}
}  // namespace vp

extern "C" {

const char* vp_last_error(void) { return g_last_error.c_str(); }
const char* vp_version(void) { return "voxphoton-mi355x 0.1 (gfx950)"; }

int vp_device_count(int* count) {
  VP_API_BEGIN
  VP_REQUIRE(count, ERR_INVALID, "count is NULL");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  VP_API_END
}

int vp_ctx_create(int device, vp_ctx* out) {
  VP_API_BEGIN
  VP_REQUIRE(out, ERR_INVALID, "out is NULL");
  int c = 0;
  VP_HIP(hipGetDeviceCount(&c));
  VP_REQUIRE(device >= 0 && device < c, ERR_INVALID, "device index out of range");
  VP_HIP(hipSetDevice(device));
  auto* ctx = new vp_ctx_s();
  ctx->device = device;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    VP_HIP(e);
  }
  register_stream(device, ctx->stream);
  *out = ctx;
  VP_API_END
}

int vp_ctx_destroy(vp_ctx ctx) {
  VP_API_BEGIN
  if (!ctx) return VP_OK;
  VP_HIP(hipSetDevice(ctx->device));
  (void)hipStreamSynchronize(ctx->stream);
  unregister_stream(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  if (ctx->scratch) dev_free(ctx->scratch);
  delete ctx;
  dev_cache_flush();   // released large buffers go back to the device
  VP_API_END
}

int vp_ctx_synchronize(vp_ctx ctx) {
  VP_API_BEGIN
  VP_REQUIRE(ctx, ERR_INVALID, "ctx is NULL");
  VP_HIP(hipSetDevice(ctx->device));
  VP_HIP(hipStreamSynchronize(ctx->stream));
  VP_API_END
}

int vp_ctx_set_recycle_width(vp_ctx ctx, int width) {
  VP_API_BEGIN
  VP_REQUIRE(ctx, ERR_INVALID, "ctx is NULL");
  VP_REQUIRE(width >= -1 && width <= 1000, ERR_INVALID, "recycle width must be -1..1000");
  ctx->recycle_width = width;
  VP_API_END
}

int vp_ctx_set_profiling(vp_ctx ctx, int on) {
  VP_API_BEGIN
  VP_REQUIRE(ctx, ERR_INVALID, "ctx is NULL");
  VP_HIP(hipSetDevice(ctx->device));
  prof_enable(ctx->stream, on != 0);
  VP_API_END
}

int vp_ctx_kernel_stats(vp_ctx ctx, const char* kernel, int64_t* launches, double* total_ms, int reset) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && kernel && launches && total_ms, ERR_INVALID, "bad arguments");
  int kid = -1;
  for (int k = 0; k < kNumKernelIds; ++k)
    if (std::string(kernel) == kernel_id_name(k)) kid = k;
  VP_REQUIRE(kid >= 0, ERR_INVALID, std::string("unknown kernel name ") + kernel);
  VP_HIP(hipSetDevice(ctx->device));
  long long l = 0;
  double ms = 0;
  prof_stats(ctx->stream, kid, &l, &ms, reset != 0);
  *launches = l;
  *total_ms = ms;
  VP_API_END
}

int vp_mem_info(vp_ctx ctx, int64_t* free_bytes, int64_t* total_bytes) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && free_bytes && total_bytes, ERR_INVALID, "bad arguments");
  VP_HIP(hipSetDevice(ctx->device));
  size_t f = 0, t = 0;
  VP_HIP(hipMemGetInfo(&f, &t));
  *free_bytes = int64_t(f);
  *total_bytes = int64_t(t);
  VP_API_END
}

int vp_malloc(vp_ctx ctx, int64_t bytes, void** dptr) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && dptr && bytes >= 0, ERR_INVALID, "vp_malloc: bad argument");
  VP_HIP(hipSetDevice(ctx->device));
  *dptr = nullptr;
  if (bytes) *dptr = dev_alloc(size_t(bytes));
  VP_API_END
}

int vp_free(vp_ctx ctx, void* dptr) {
  VP_API_BEGIN
  VP_REQUIRE(ctx, ERR_INVALID, "ctx is NULL");
  VP_HIP(hipSetDevice(ctx->device));
  if (dptr) dev_free(dptr);
  VP_API_END
}

int vp_memcpy_h2d(vp_ctx ctx, void* dst, const void* src, int64_t bytes) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && dst && src && bytes >= 0, ERR_INVALID, "vp_memcpy_h2d: bad argument");
  VP_HIP(hipSetDevice(ctx->device));
  VP_HIP(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyHostToDevice, ctx->stream));
  VP_HIP(hipStreamSynchronize(ctx->stream));
  VP_API_END
}

int vp_memcpy_d2h(vp_ctx ctx, void* dst, const void* src, int64_t bytes) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && dst && src && bytes >= 0, ERR_INVALID, "vp_memcpy_d2h: bad argument");
  VP_HIP(hipSetDevice(ctx->device));
  VP_HIP(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDeviceToHost, ctx->stream));
  VP_HIP(hipStreamSynchronize(ctx->stream));
  VP_API_END
}

}  // extern "C"

// ----------------------------------------------------------------------------
// Operator construction
// ----------------------------------------------------------------------------
namespace vp {

// Validates idx (1-based into the 3V field) and builds the voxel map: for the
// [idxS; V+idxS; 2V+idxS] form with sorted idxS a V-entry voxel -> position
// map (structured), otherwise a 3V-entry field -> position map.
void index_map(int64_t V, const int64_t* idx, int64_t n, std::vector<int>& pos, int& structured_out,
               int64_t& nS_out) {
  VP_REQUIRE(idx && n > 0, ERR_INVALID, "idx must hold at least one index");
  VP_REQUIRE(n <= 3 * V, ERR_INVALID, "more indices than field entries");
  VP_REQUIRE(3 * V < (int64_t(1) << 31), ERR_UNSUPPORTED, "grid too large for 32-bit voxel maps");
  for (int64_t i = 0; i < n; ++i)
    VP_REQUIRE(idx[i] >= 1 && idx[i] <= 3 * V, ERR_INVALID,
               "idx[" + std::to_string(i) + "] outside 1..3*L*M*N");
  bool structured = (n % 3 == 0);
  const int64_t nS = n / 3;
  if (structured) {
    for (int64_t i = 0; i < nS && structured; ++i) {
      if (idx[i] > V) structured = false;
      else if (i > 0 && idx[i] <= idx[i - 1]) structured = false;
      else if (idx[nS + i] != idx[i] + V || idx[2 * nS + i] != idx[i] + 2 * V) structured = false;
    }
  }
  if (structured) {
    structured_out = 1;
    nS_out = nS;
    pos.assign(size_t(V), -1);
    for (int64_t i = 0; i < nS; ++i) pos[size_t(idx[i] - 1)] = int(i);
  } else {
    structured_out = 0;
    nS_out = n;
    pos.assign(size_t(3 * V), -1);
    for (int64_t i = 0; i < n; ++i) {
      VP_REQUIRE(pos[size_t(idx[i] - 1)] < 0, ERR_INVALID, "duplicate index in idx");
      pos[size_t(idx[i] - 1)] = int(i);
    }
  }
}

// Validates idx and fills the voxel map of an operator.
static void index_setup(vp_op_s* op, const int64_t* idx, int64_t n, std::vector<int>& pos) {
  index_map(op->V, idx, n, pos, op->structured, op->nS);
  op->n = n;
}

// Bounding box of the voxels that carry unknowns.  The passes run on the box
// only; VP_MATVEC_PRUNE=0 keeps the full grid.
static void unknown_box(vp_op_s* op, const std::vector<int>& pos) {
  const char* e = std::getenv("VP_MATVEC_PRUNE");
  const bool prune = !(e && std::strcmp(e, "0") == 0);
  int lo[3] = {op->L, op->M, op->N}, hi[3] = {-1, -1, -1};
  if (prune) {
    for (size_t i = 0; i < pos.size(); ++i) {
      if (pos[i] < 0) continue;
      int64_t v = int64_t(i) % op->V;
      const int c[3] = {int(v % op->L), int((v / op->L) % op->M), int(v / (int64_t(op->L) * op->M))};
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], c[a]);
        hi[a] = std::max(hi[a], c[a]);
      }
    }
  }
  if (!prune || hi[0] < 0) {
    lo[0] = lo[1] = lo[2] = 0;
    hi[0] = op->L - 1; hi[1] = op->M - 1; hi[2] = op->N - 1;
  }
  op->l0 = lo[0]; op->m0 = lo[1]; op->n0 = lo[2];
  op->Lb = hi[0] - lo[0] + 1; op->Mb = hi[1] - lo[1] + 1; op->Nb = hi[2] - lo[2] + 1;
}

// Gathers a complex128 grid array (L x M x N) to the unknown entries.
static std::vector<double2> gather_material(const vp_op_s* op, const double* A, const int64_t* idx, double dflt) {
  const int64_t cnt = op->structured ? op->nS : op->n;
  std::vector<double2> out(static_cast<size_t>(cnt));
  for (int64_t i = 0; i < cnt; ++i) {
    const int64_t v = (idx[i] - 1) % op->V;
    out[size_t(i)] = A ? make_double2(A[2 * v], A[2 * v + 1]) : make_double2(dflt, 0.0);
  }
  return out;
}

template <typename T>
static void init_prec(vp_op_s* op, OpPrec<T>& P, bool fp64, const std::vector<double2>& mc,
                      const std::vector<double2>* mr, double gram_t) {
  (void)fp64;
  P.present = true;
  const size_t cnt = mc.size();
  std::vector<C<T>> h(cnt);
  for (size_t i = 0; i < cnt; ++i) h[i] = mk<T>(T(mc[i].x), T(mc[i].y));
  P.mc.alloc(cnt * sizeof(C<T>));
  VP_HIP(hipMemcpy(P.mc.p, h.data(), cnt * sizeof(C<T>), hipMemcpyHostToDevice));
  if (mr) {
    for (size_t i = 0; i < cnt; ++i) h[i] = mk<T>(T((*mr)[i].x), T((*mr)[i].y));
    P.mr.alloc(cnt * sizeof(C<T>));
    VP_HIP(hipMemcpy(P.mr.p, h.data(), cnt * sizeof(C<T>), hipMemcpyHostToDevice));
  }
  P.gram = T(gram_t);
  P.xs.alloc(size_t(op->n) * sizeof(C<T>));
  P.ys.alloc(size_t(op->n) * sizeof(C<T>));
}

template <typename T>
static void upload_probs(vp_op_s* op, OpPrec<T>& P) {
  ProbPtrs<T> pp{};
  pp.work = P.work.template as<C<T>>();
  pp.work2 = P.work2.p ? P.work2.template as<C<T>>() : nullptr;
  pp.fN = P.fN.template as<C<T>>();
  pp.pos = op->pos.as<int>();
  pp.mc = P.mc.template as<C<T>>();
  pp.mr = P.mr.p ? P.mr.template as<C<T>>() : nullptr;
  pp.gram = P.gram;
  P.probs.alloc(sizeof(pp));
  VP_HIP(hipMemcpy(P.probs.p, &pp, sizeof(pp), hipMemcpyHostToDevice));
}

// Common part of every vp_op_create*: everything except the tensor.
static vp_op_s* op_begin(vp_ctx ctx, int64_t L, int64_t M, int64_t N, int64_t Lp, int64_t Mp, int64_t Np,
                         const double* Mr, const double* Mc, double gram, const int64_t* idx, int64_t n,
                         int precisions) {
  VP_REQUIRE(ctx, ERR_INVALID, "ctx is NULL");
  VP_REQUIRE(L > 0 && M > 0 && N > 0, ERR_INVALID, "grid extents must be positive");
  VP_REQUIRE(Lp >= 0 && Mp >= 0 && Np >= 0, ERR_INVALID, "FFT sizes must be >= 0 (0: chosen by the autotuner)");
  VP_REQUIRE(Lp < (1 << 20) && Mp < (1 << 20) && Np < (1 << 20), ERR_UNSUPPORTED, "FFT size too large");
  VP_REQUIRE(Mc, ERR_INVALID, "Mc is NULL");
  VP_REQUIRE(precisions & (VP_FP64 | VP_FP32), ERR_INVALID, "precisions must include VP_FP64 or VP_FP32");
  VP_REQUIRE(std::isfinite(gram) && gram > 0, ERR_INVALID, "gram must be positive");
  VP_HIP(hipSetDevice(ctx->device));
  auto* op = new vp_op_s();
  try {
    op->ctx = ctx;
    op->L = int(L); op->M = int(M); op->N = int(N);
    op->Lp = int(Lp); op->Mp = int(Mp); op->Np = int(Np);
    op->V = L * M * N;
    op->gram = gram;
    op->scale32 = gram;
    op->has_mr = Mr != nullptr;
    std::vector<int> pos;
    index_setup(op, idx, n, pos);
    unknown_box(op, pos);
    // only offsets inside the unknowns' box reach the result: any circulant
    // of at least 2B-1 per axis gives the same operator on idx
    VP_REQUIRE((Lp == 0 || Lp >= 2 * op->Lb - 1) && (Mp == 0 || Mp >= 2 * op->Mb - 1) &&
                   (Np == 0 || Np >= 2 * op->Nb - 1),
               ERR_INVALID,
               "FFT sizes must be at least 2*B-1 for the unknowns' bounding box B (" + std::to_string(op->Lb) +
                   " x " + std::to_string(op->Mb) + " x " + std::to_string(op->Nb) + ")");
    op->pos.alloc(pos.size() * sizeof(int));
    VP_HIP(hipMemcpy(op->pos.p, pos.data(), pos.size() * sizeof(int), hipMemcpyHostToDevice));
    auto mc = gather_material(op, Mc, idx, 0.0);
    std::vector<double2> mr;
    if (Mr) mr = gather_material(op, Mr, idx, 1.0);
    op->unit_contrast = !Mr && std::all_of(mc.begin(), mc.end(), [](double2 v) { return v.x == 1.0 && v.y == 0.0; });
    op->all_voxels = n == 3 * op->V;
    for (int64_t i = 0; op->all_voxels && i < n; ++i) op->all_voxels = idx[i] == i + 1;
    if (precisions & VP_FP64) init_prec(op, op->p64, true, mc, Mr ? &mr : nullptr, gram);
    if (precisions & VP_FP32) init_prec(op, op->p32, false, mc, Mr ? &mr : nullptr, gram / op->scale32);
  } catch (...) {
    delete op;
    throw;
  }
  return op;
}

// Matvec variant: VP_MATVEC_VARIANT = fused | split | auto (default).
// -1 means "time both".
static int variant_request() {
  const char* e = std::getenv("VP_MATVEC_VARIANT");
  if (!e || !*e || !std::strcmp(e, "auto")) return -1;
  if (!std::strcmp(e, "fused")) return 0;
  if (!std::strcmp(e, "split")) return 1;
  throw VpError{ERR_INVALID, "VP_MATVEC_VARIANT must be fused, split or auto"};
}

// FFT sizes of one precision's plan.
struct Sizes {
  int Lp, Mp, Np;
};

// Geometry, twiddles, tensor layout and workspaces of one precision for one
// matvec variant at FFT sizes z.  F is the true-frequency FP64 tensor at
// those sizes, or nullptr for a zero tensor of the right layout (timing).
template <typename T>
static void configure_prec(vp_op_s* op, OpPrec<T>& P, bool fp64, const double2* F, Sizes z, double scale,
                           int variant, bool fsym) {
  hipStream_t s = op->ctx->stream;
  build_geometry(P.g, P.plan, *op, z.Lp, z.Mp, z.Np, fp64, op->ctx->device, variant);
  upload_twiddles(P, s);
  P.plan.fsym = (variant == 1 && fsym) ? 1 : 0;
  if (F) {
    if (P.plan.fsym)
      layout_fN_sym<T>(P.g, F, P.fN, scale, s);
    else
      layout_fN<T>(P.g, F, P.fN, scale, s);
  } else {
    const int64_t cnt = P.plan.fsym ? int64_t(z.Lp / 2 + 1) * (z.Mp / 2 + 1) * (z.Np / 2 + 1) * 6
                                    : int64_t(z.Lp) * z.Mp * z.Np * 6;
    if (P.fN.bytes != size_t(cnt) * sizeof(C<T>)) P.fN.alloc(size_t(cnt) * sizeof(C<T>));
    VP_HIP(hipMemsetAsync(P.fN.p, 0, P.fN.bytes, s));
  }
  P.work.alloc(size_t(3) * z.Lp * op->Mb * op->Nb * sizeof(C<T>));
  if (variant == 1)
    P.work2.alloc(size_t(3) * z.Lp * z.Mp * op->Nb * sizeof(C<T>));
  else
    P.work2.release();
  upload_probs(op, P);
}

// Mean time of a few applies on zero input, in ms.
template <typename T>
static double time_applies(vp_op_s* op, OpPrec<T>& P) {
  hipStream_t s = op->ctx->stream;
  VP_HIP(hipMemsetAsync(P.xs.p, 0, P.xs.bytes, s));
  C<T>* x = P.xs.template as<C<T>>();
  C<T>* y = P.ys.template as<C<T>>();
  op_apply_dev<T>(op, 0, x, y);
  hipEvent_t a, b;
  VP_HIP(hipEventCreate(&a));
  VP_HIP(hipEventCreate(&b));
  const int reps = 3;
  float ms = 0;
  hipError_t err = hipEventRecord(a, s);
  for (int i = 0; i < reps && err == hipSuccess; ++i) op_apply_dev<T>(op, 0, x, y);
  if (err == hipSuccess) err = hipEventRecord(b, s);
  if (err == hipSuccess) err = hipEventSynchronize(b);
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  VP_HIP(err);
  return double(ms) / reps;
}

// Timed choice between the fused and split matvec, made once per
// (device, unknowns' box, FFT sizes, precision) in a process so that operators of one
// shape always share a plan (batched solves rely on it).
struct VariantKey {
  int device, Lb, Mb, Nb, Lp, Mp, Np, fp64;
  bool operator<(const VariantKey& o) const {
    return std::tie(device, Lb, Mb, Nb, Lp, Mp, Np, fp64) <
           std::tie(o.device, o.Lb, o.Mb, o.Nb, o.Lp, o.Mp, o.Np, o.fp64);
  }
};
static std::mutex g_variant_mu;
static std::map<VariantKey, int> g_variant;

template <typename T>
static void finalize_prec(vp_op_s* op, OpPrec<T>& P, bool fp64, const double2* F, Sizes z, double scale,
                          bool fsym) {
  const VariantKey key{op->ctx->device, op->Lb, op->Mb, op->Nb, z.Lp, z.Mp, z.Np, int(fp64)};
  int v = variant_request();
  if (v < 0) {
    std::lock_guard<std::mutex> lk(g_variant_mu);
    auto it = g_variant.find(key);
    if (it != g_variant.end()) v = it->second;
  }
  if (v >= 0) {
    configure_prec(op, P, fp64, F, z, scale, v, fsym);
    return;
  }
  double best = 0;
  int best_v = -1, configured = -1;
  std::string why;
  for (int cand : {0, 1}) {
    try {
      configure_prec(op, P, fp64, F, z, scale, cand, fsym);
    } catch (const VpError& e) {
      if (e.code != ERR_UNSUPPORTED) throw;
      why = e.msg;
      configured = -1;
      continue;
    }
    configured = cand;
    const double t = time_applies(op, P);
    if (best_v < 0 || t < best) { best = t; best_v = cand; }
  }
  VP_REQUIRE(best_v >= 0, ERR_UNSUPPORTED, why);
  if (configured != best_v) configure_prec(op, P, fp64, F, z, scale, best_v, fsym);
  std::lock_guard<std::mutex> lk(g_variant_mu);
  g_variant.emplace(key, best_v);
}

// ---------------------------------------------------------------------------
// FFT-size autotuner.  Any circulant of at least 2B-1 points per axis (B the
// unknowns' box) gives the same operator, so the sizes are free.  Per axis
// the candidates are the lengths in [2B-1, 2B-1 + max(8, (2B-1)/4)] whose
// prime factors the device FFTs handle; lengths with compile-time kernels
// (x: class length Lp/2, y: the y pass, z: a fused radix pair) come first,
// two of them at most, else the two shortest.  Every combination is timed
// with the split matvec on a zero tensor (the kernels' time does not depend
// on the values), and the fastest is kept per (device, box, precision) for
// the process, so operators of one shape share their sizes and batch
// together.  VP_FFT_SIZES=reference selects the reference's 2L x 2M x 2N,
// VP_FFT_SIZES=min the shortest candidates without timing.
// ---------------------------------------------------------------------------
static bool smooth_length(int n, bool fp64) {
  for (int q : {2, 3, 5, 7, 11, 13}) {
    if (q == 13 && fp64) break;
    while (n % q == 0) n /= q;
  }
  return n == 1;
}

// A length with a compile-time kernel on its axis (x: the class length
// n/2, y: the y pass, z: a fused radix pair).
static bool fast_length(int n, int axis, bool fp64) {
  int ra = 0, rb = 0;
  return axis == 0 ? (n % 2 == 0 && xct_threads(n / 2, fp64) > 0)
       : axis == 1 ? yct_supported(n)
                   : zfused_pair(n, fp64, ra, rb);
}

static std::vector<int> axis_candidates(int B, int axis, bool fp64, int fixed) {
  if (fixed > 0) return {fixed};
  const int lo = std::max(2, 2 * B - 1);
  const int hi = lo + std::max(8, lo / 4);
  std::vector<int> fast, slow;
  for (int n = lo; n <= hi; ++n) {
    if (!smooth_length(n, fp64)) continue;
    (fast_length(n, axis, fp64) ? fast : slow).push_back(n);
  }
  std::vector<int> out(fast.begin(), fast.begin() + std::min<size_t>(2, fast.size()));
  for (size_t i = 0; out.size() < 2 && i < slow.size(); ++i)
    if (out.empty() || slow[i] < out[0] || fast.empty()) out.push_back(slow[i]);
  std::sort(out.begin(), out.end());
  return out;
}

struct SizeKey {
  int device, Lb, Mb, Nb, fp64, fsym;
  int Lp, Mp, Np;   // the caller's fixed sizes (0 free)
  int La, Ma, Na;   // the sizes of a caller's tensor, candidates besides the smooth lengths (0 none)
  bool operator<(const SizeKey& o) const {
    return std::tie(device, Lb, Mb, Nb, fp64, fsym, Lp, Mp, Np, La, Ma, Na) <
           std::tie(o.device, o.Lb, o.Mb, o.Nb, o.fp64, o.fsym, o.Lp, o.Mp, o.Np, o.La, o.Ma, o.Na);
  }
};
static std::mutex g_sizes_mu;
static std::map<SizeKey, Sizes> g_sizes;

// 0 auto (timed), 1 reference 2D, 2 shortest candidates, 3 per axis the
// shortest candidate with a compile-time kernel, else the shortest (no timing)
static int fft_sizes_mode() {
  const char* e = std::getenv("VP_FFT_SIZES");
  if (!e || !*e || !std::strcmp(e, "auto")) return 0;
  if (!std::strcmp(e, "reference")) return 1;
  if (!std::strcmp(e, "min")) return 2;
  if (!std::strcmp(e, "kernels")) return 3;
  throw VpError{ERR_INVALID, "VP_FFT_SIZES must be auto, reference, min or kernels"};
}

// Circulant sizes for one precision: fixed axes (> 0) are kept; each free
// axis takes the faster of its smooth candidates (and of the caller's
// tensor size `also`, when given), timed on the device.
template <typename T>
static Sizes choose_sizes(vp_op_s* op, OpPrec<T>& P, bool fp64, bool fsym, Sizes fixed, Sizes also) {
  const int mode = fft_sizes_mode();
  if (mode == 1) {
    auto pick = [](int f, int a, int D) { return f ? f : a ? a : 2 * D; };
    return Sizes{pick(fixed.Lp, also.Lp, op->L), pick(fixed.Mp, also.Mp, op->M), pick(fixed.Np, also.Np, op->N)};
  }
  auto cands = [&](int B, int axis, int f, int a) {
    auto c = axis_candidates(B, axis, fp64, f);
    if (!f && a > 0 && std::find(c.begin(), c.end(), a) == c.end()) {
      c.push_back(a);
      std::sort(c.begin(), c.end());
    }
    return c;
  };
  const auto cx = cands(op->Lb, 0, fixed.Lp, also.Lp), cy = cands(op->Mb, 1, fixed.Mp, also.Mp),
             cz = cands(op->Nb, 2, fixed.Np, also.Np);
  VP_REQUIRE(!cx.empty() && !cy.empty() && !cz.empty(), ERR_UNSUPPORTED, "no FFT size candidates for the box");
  if (mode == 2 || (cx.size() == 1 && cy.size() == 1 && cz.size() == 1)) return Sizes{cx[0], cy[0], cz[0]};
  if (mode == 3) {
    auto pick = [&](const std::vector<int>& c, int axis) {
      for (int n : c)
        if (fast_length(n, axis, fp64)) return n;
      return c[0];
    };
    return Sizes{pick(cx, 0), pick(cy, 1), pick(cz, 2)};
  }
  const SizeKey key{op->ctx->device, op->Lb, op->Mb, op->Nb, int(fp64), int(fsym), fixed.Lp, fixed.Mp, fixed.Np,
                    also.Lp, also.Mp, also.Np};
  {
    std::lock_guard<std::mutex> lk(g_sizes_mu);
    auto it = g_sizes.find(key);
    if (it != g_sizes.end()) return it->second;
  }
  Sizes best{0, 0, 0};
  double tbest = 0;
  std::string why = "no FFT size candidate plans";
  for (int a : cx)
    for (int b : cy)
      for (int c : cz) {
        const Sizes z{a, b, c};
        try {
          configure_prec(op, P, fp64, nullptr, z, 1.0, 1, fsym);
        } catch (const VpError& e) {
          if (e.code != ERR_UNSUPPORTED) throw;
          why = e.msg;
          continue;
        }
        const double t = time_applies(op, P);
        if (best.Lp == 0 || t < tbest) { best = z; tbest = t; }
      }
  if (best.Lp == 0) {   // the split passes do not plan at any candidate: shortest sizes, any variant
    best = Sizes{cx[0], cy[0], cz[0]};
  }
  std::lock_guard<std::mutex> lk(g_sizes_mu);
  g_sizes.emplace(key, best);
  return best;
}

// Source of an operator's tensor: the caller's true-frequency tensor F at
// its sizes `given` (kept, or resampled to the sizes a precision's plan
// chooses), or a device Toeplitz tensor on the unknowns' box
// (Lb x Mb x Nb x 6) embedded at the chosen sizes.
struct TensorSource {
  const double2* F = nullptr;
  Sizes given{0, 0, 0};
  const double2* Gbox = nullptr;
};

// VP_FN_RESAMPLE=0 keeps a caller's tensor at its own sizes.
static bool resample_allowed() {
  const char* e = std::getenv("VP_FN_RESAMPLE");
  return !(e && std::strcmp(e, "0") == 0);
}

// Common tail of every vp_op_create*.
static void op_finish(vp_op_s* op, const TensorSource& src) {
  hipStream_t s = op->ctx->stream;
  // VP_FN_SYMMETRY=0 keeps the full tensor in every variant.
  const char* e = std::getenv("VP_FN_SYMMETRY");
  const bool sym_allowed = !(e && std::strcmp(e, "0") == 0);
  auto same = [](Sizes a, Sizes b) { return a.Lp == b.Lp && a.Mp == b.Mp && a.Np == b.Np; };
  DevBuf built;            // tensor embedded from Gbox, or resampled from F, at built_z
  Sizes built_z{0, 0, 0};
  // the tensor at sizes z; for a caller's tensor z may fall back to its own
  // sizes when a length does not resample
  auto tensor_at = [&](Sizes& z) -> const double2* {
    if (src.F && same(z, src.given)) return src.F;
    if (built.p && same(built_z, z)) return built.as<double2>();
    if (src.F) {
      const int P0[3] = {src.given.Lp, src.given.Mp, src.given.Np}, B[3] = {op->Lb, op->Mb, op->Nb},
                P1[3] = {z.Lp, z.Mp, z.Np};
      DevBuf r;
      if (!resample_fN(src.F, P0, B, P1, op->ctx->device, r, s)) {
        z = src.given;
        return src.F;
      }
      built = std::move(r);
    } else {
      built.release();
      built.alloc(size_t(z.Lp) * z.Mp * z.Np * 6 * sizeof(double2));
      toeplitz_to_fN(op->Lb, op->Mb, op->Nb, z.Lp, z.Mp, z.Np, op->ctx->device, src.Gbox, built.as<double2>(), s);
    }
    built_z = z;
    return built.as<double2>();
  };
  // symmetry: a caller's tensor is checked; an embedded or resampled tensor
  // is parity symmetric by construction (checked all the same)
  bool fsym = sym_allowed;
  if (src.F) fsym = sym_allowed && fN_parity_symmetric(src.given.Lp, src.given.Mp, src.given.Np, src.F, s);
  const bool resample = src.F && resample_allowed();
  const Sizes fixed = src.F && !resample ? src.given : Sizes{op->Lp, op->Mp, op->Np};
  const Sizes also = resample ? src.given : Sizes{0, 0, 0};
  auto one = [&](auto tag, bool fp64, double scale_div) {
    using T = decltype(tag);
    OpPrec<T>& P = prec_of<T>(op);
    if (!P.present) return;
    Sizes z = choose_sizes<T>(op, P, fp64, fsym, fixed, also);
    const double2* F = tensor_at(z);
    if (F != src.F && fsym) fsym = fN_parity_symmetric(z.Lp, z.Mp, z.Np, F, s);
    const double Pn = double(z.Lp) * z.Mp * z.Np;
    finalize_prec(op, P, fp64, F, z, 1.0 / (Pn * scale_div), fsym);
  };
  one(double{}, true, 1.0);
  one(float{}, false, op->scale32);
  op->fsym = fsym;
  VP_HIP(hipStreamSynchronize(s));
}

template <typename T>
void op_apply_dev(vp_op_s* op, int transp, const C<T>* x, C<T>* y) {
  OpPrec<T>& P = prec_of<T>(op);
  VP_REQUIRE(P.present, ERR_INVALID, "operator was not created for this precision");
  VecArgs<T> va{x, y, 0, 0, nullptr};
  launch_matvec<T>(P.g, P.plan, P.probs.template as<ProbPtrs<T>>(), va, 1, transp, P.twx(), P.twy(), P.twz(),
                   op->ctx->stream);
}
template void op_apply_dev<double>(vp_op_s*, int, const double2*, double2*);
template void op_apply_dev<float>(vp_op_s*, int, const float2*, float2*);

}  // namespace vp

namespace {
// every operator of a combination shares the geometry, scaling, precisions
// and plan of ops[0], so that their stored tensors combine entry by entry
void require_combinable(const vp_op_s* o0, const vp_op_s* o) {
  auto same_sizes = [](const Geometry& a, const Geometry& b) {
    return a.x.P == b.x.P && a.y.P == b.y.P && a.z.P == b.z.P;
  };
  VP_REQUIRE(o && o->ctx == o0->ctx && o->L == o0->L && o->M == o0->M && o->N == o0->N &&
                 same_sizes(o->p64.g, o0->p64.g) && same_sizes(o->p32.g, o0->p32.g) && o->n == o0->n &&
                 o->gram == o0->gram &&
                 o->scale32 == o0->scale32 && o->fsym == o0->fsym && o->p64.present == o0->p64.present &&
                 o->p32.present == o0->p32.present && o->p64.fN.bytes == o0->p64.fN.bytes &&
                 o->p32.fN.bytes == o0->p32.fN.bytes && o->p64.plan.variant == o0->p64.plan.variant &&
                 o->p32.plan.variant == o0->p32.plan.variant,
             ERR_INVALID, "combined operators must share geometry, scaling and plan");
}

// out.fN = sum_i w[i] ops[i].fN per present precision (stream-ordered)
void combine_tensors(vp_op_s* out, const vp_op* ops, const double* w, int k, hipStream_t s) {
  std::vector<double2> wv(k);
  for (int i = 0; i < k; ++i) wv[i] = make_double2(w[2 * i], w[2 * i + 1]);
  auto one = [&](auto tag) {
    using T = decltype(tag);
    OpPrec<T>& P = prec_of<T>(out);
    if (!P.present) return;
    std::vector<const C<T>*> src(k);
    for (int i = 0; i < k; ++i) src[i] = prec_of<T>(const_cast<vp_op_s*>(ops[i])).fN.template as<C<T>>();
    tensor_lincomb<T>(P.fN.template as<C<T>>(), src.data(), wv.data(), k, int64_t(P.fN.bytes / sizeof(C<T>)), s);
  };
  one(double{});
  one(float{});
}
}  // namespace

extern "C" {

int vp_op_create(vp_ctx ctx, int64_t L, int64_t M, int64_t N, int64_t Lp, int64_t Mp, int64_t Np, const double* fN,
                 const double* Mr, const double* Mc, double gram, const int64_t* idx, int64_t n, int precisions,
                 vp_op* out) {
  VP_API_BEGIN
  VP_REQUIRE(out && fN, ERR_INVALID, "out/fN is NULL");
  VP_REQUIRE(Lp > 0 && Mp > 0 && Np > 0, ERR_INVALID, "fN extents must be given");
  vp_op_s* op = op_begin(ctx, L, M, N, Lp, Mp, Np, Mr, Mc, gram, idx, n, precisions);
  try {
    const size_t bytes = size_t(op->Lp) * op->Mp * op->Np * 6 * sizeof(double2);
    DevBuf F(bytes);
    VP_HIP(hipMemcpy(F.p, fN, bytes, hipMemcpyHostToDevice));
    TensorSource src;
    src.F = F.as<double2>();
    src.given = Sizes{op->Lp, op->Mp, op->Np};
    op->Lp = op->Mp = op->Np = 0;   // the sizes are free: the tensor resamples
    op_finish(op, src);
  } catch (...) {
    delete op;
    throw;
  }
  *out = op;
  VP_API_END
}

int vp_op_create_toeplitz(vp_ctx ctx, int64_t L, int64_t M, int64_t N, int64_t Lp, int64_t Mp, int64_t Np,
                          const double* G, const double* Mr, const double* Mc, double gram, const int64_t* idx,
                          int64_t n, int precisions, vp_op* out) {
  VP_API_BEGIN
  VP_REQUIRE(out && G, ERR_INVALID, "out/G is NULL");
  vp_op_s* op = op_begin(ctx, L, M, N, Lp, Mp, Np, Mr, Mc, gram, idx, n, precisions);
  try {
    // the offsets inside the unknowns' box: G(0:Lb, 0:Mb, 0:Nb, :)
    const int64_t Vb = int64_t(op->Lb) * op->Mb * op->Nb;
    std::vector<double2> hb(size_t(Vb) * 6);
    for (int k = 0; k < 6; ++k)
      for (int c = 0; c < op->Nb; ++c)
        for (int b = 0; b < op->Mb; ++b) {
          const double* src = G + 2 * (int64_t(op->L) * (b + int64_t(op->M) * (c + int64_t(op->N) * k)));
          double2* dst = hb.data() + int64_t(op->Lb) * (b + int64_t(op->Mb) * (c + int64_t(op->Nb) * k));
          std::memcpy(dst, src, size_t(op->Lb) * sizeof(double2));
        }
    DevBuf dG(hb.size() * sizeof(double2));
    VP_HIP(hipMemcpy(dG.p, hb.data(), dG.bytes, hipMemcpyHostToDevice));
    TensorSource src;
    src.Gbox = dG.as<double2>();
    op_finish(op, src);
  } catch (...) {
    delete op;
    throw;
  }
  *out = op;
  VP_API_END
}

int vp_op_create_assembled(vp_ctx ctx, int64_t L, int64_t M, int64_t N, int64_t Lp, int64_t Mp, int64_t Np,
                           double dx, double ko, const double* Mr, const double* Mc, const int64_t* idx, int64_t n,
                           int precisions, vp_op* out) {
  VP_API_BEGIN
  VP_REQUIRE(out, ERR_INVALID, "out is NULL");
  VP_REQUIRE(dx > 0 && std::isfinite(ko) && ko >= 0, ERR_INVALID, "dx must be positive and ko finite");
  vp_op_s* op = op_begin(ctx, L, M, N, Lp, Mp, Np, Mr, Mc, dx * dx * dx, idx, n, precisions);
  try {
    // the Green tensor at the offsets inside the unknowns' box only
    DevBuf dG(size_t(op->Lb) * op->Mb * op->Nb * 6 * sizeof(double2));
    assemble_toeplitz_dev(op->Lb, op->Mb, op->Nb, dx, ko, dG.as<double2>(), ctx->stream);
    TensorSource src;
    src.Gbox = dG.as<double2>();
    op_finish(op, src);
  } catch (...) {
    delete op;
    throw;
  }
  *out = op;
  VP_API_END
}

int vp_assemble_toeplitz(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko, double* G, int mem) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && G, ERR_INVALID, "ctx/G is NULL");
  VP_REQUIRE(L > 0 && M > 0 && N > 0 && dx > 0, ERR_INVALID, "bad grid");
  VP_HIP(hipSetDevice(ctx->device));
  const size_t bytes = size_t(L * M * N) * 6 * sizeof(double2);
  if (mem == VP_MEM_DEVICE) {
    assemble_toeplitz_dev(int(L), int(M), int(N), dx, ko, reinterpret_cast<double2*>(G), ctx->stream);
  } else {
    DevBuf dG(bytes);
    assemble_toeplitz_dev(int(L), int(M), int(N), dx, ko, dG.as<double2>(), ctx->stream);
    VP_HIP(hipMemcpy(G, dG.p, bytes, hipMemcpyDeviceToHost));
  }
  VP_API_END
}

int vp_op_apply(vp_op op, int transp, const void* x, void* y, int mem, int prec) {
  VP_API_BEGIN
  VP_REQUIRE(op && x && y, ERR_INVALID, "op/x/y is NULL");
  VP_REQUIRE(x != y, ERR_INVALID, "x and y must not alias");
  VP_REQUIRE(transp == VP_NOTRANSP || transp == VP_TRANSP, ERR_INVALID, "transp must be 0 or 1");
  VP_REQUIRE(prec == VP_FP64 || prec == VP_FP32, ERR_INVALID, "prec must be VP_FP64 or VP_FP32");
  VP_HIP(hipSetDevice(op->ctx->device));
  hipStream_t s = op->ctx->stream;
  const int64_t n = op->n;
  if (mem == VP_MEM_DEVICE) {
    if (prec == VP_FP64) op_apply_dev<double>(op, transp, static_cast<const double2*>(x), static_cast<double2*>(y));
    else op_apply_dev<float>(op, transp, static_cast<const float2*>(x), static_cast<float2*>(y));
  } else if (mem == VP_MEM_HOST) {
    if (prec == VP_FP64) {
      VP_REQUIRE(op->p64.present, ERR_INVALID, "operator has no FP64 copy");
      VP_HIP(hipMemcpyAsync(op->p64.xs.p, x, n * sizeof(double2), hipMemcpyHostToDevice, s));
      op_apply_dev<double>(op, transp, op->p64.xs.as<double2>(), op->p64.ys.as<double2>());
      VP_HIP(hipMemcpyAsync(y, op->p64.ys.p, n * sizeof(double2), hipMemcpyDeviceToHost, s));
    } else {
      VP_REQUIRE(op->p32.present, ERR_INVALID, "operator has no FP32 copy");
      // stage complex128 through the FP64 staging area of the FP32 copy
      DevBuf t(n * sizeof(double2));
      VP_HIP(hipMemcpyAsync(t.p, x, n * sizeof(double2), hipMemcpyHostToDevice, s));
      convert_c128_to_c64(t.as<double2>(), op->p32.xs.as<float2>(), n, 1.0, s);
      op_apply_dev<float>(op, transp, op->p32.xs.as<float2>(), op->p32.ys.as<float2>());
      convert_c64_to_c128(op->p32.ys.as<float2>(), t.as<double2>(), n, op->scale32, s);
      VP_HIP(hipMemcpyAsync(y, t.p, n * sizeof(double2), hipMemcpyDeviceToHost, s));
      VP_HIP(hipStreamSynchronize(s));
    }
    VP_HIP(hipStreamSynchronize(s));
  } else {
    throw VpError{ERR_INVALID, "mem must be VP_MEM_HOST or VP_MEM_DEVICE"};
  }
  VP_API_END
}

int vp_op_apply_batched(const vp_op* ops, int nvec, int transp, const void* x, void* y, int prec) {
  VP_API_BEGIN
  VP_REQUIRE(ops && nvec > 0 && x && y, ERR_INVALID, "bad batched arguments");
  VP_REQUIRE(prec == VP_FP64 || prec == VP_FP32, ERR_INVALID, "prec must be VP_FP64 or VP_FP32");
  vp_op_s* o0 = ops[0];
  VP_REQUIRE(o0, ERR_INVALID, "ops[0] is NULL");
  for (int i = 1; i < nvec; ++i) {
    VP_REQUIRE(ops[i] && ops[i]->L == o0->L && ops[i]->M == o0->M && ops[i]->N == o0->N &&
                   ops[i]->n == o0->n && ops[i]->structured == o0->structured && ops[i]->l0 == o0->l0 &&
                   ops[i]->m0 == o0->m0 && ops[i]->n0 == o0->n0 && ops[i]->Lb == o0->Lb && ops[i]->Mb == o0->Mb &&
                   ops[i]->Nb == o0->Nb,
               ERR_INVALID, "batched operators must share one geometry");
  }
  VP_HIP(hipSetDevice(o0->ctx->device));
  hipStream_t s = o0->ctx->stream;
  auto run = [&](auto tag) {
    using T = decltype(tag);
    std::vector<ProbPtrs<T>> h(nvec);
    for (int i = 0; i < nvec; ++i) {
      OpPrec<T>& P = prec_of<T>(ops[i]);
      VP_REQUIRE(P.present, ERR_INVALID, "operator lacks the requested precision");
      VP_HIP(hipMemcpy(&h[i], P.probs.p, sizeof(ProbPtrs<T>), hipMemcpyDeviceToHost));
    }
    OpPrec<T>& P0 = prec_of<T>(o0);
    // all problems must have identical class splits; rebuild if a later op chose differently
    for (int i = 1; i < nvec; ++i) {
      OpPrec<T>& P = prec_of<T>(ops[i]);
      VP_REQUIRE(P.g.x.P == P0.g.x.P && P.g.y.P == P0.g.y.P && P.g.z.P == P0.g.z.P && P.g.y.R == P0.g.y.R &&
                     P.g.z.R == P0.g.z.R && P.plan.variant == P0.plan.variant && P.plan.fsym == P0.plan.fsym,
                 ERR_INVALID, "batched operators differ in plan");
    }
    DevBuf d(sizeof(ProbPtrs<T>) * nvec);
    VP_HIP(hipMemcpyAsync(d.p, h.data(), d.bytes, hipMemcpyHostToDevice, s));
    VecArgs<T> va{static_cast<const C<T>*>(x), static_cast<C<T>*>(y), o0->n, o0->n, nullptr};
    launch_matvec<T>(P0.g, P0.plan, d.as<ProbPtrs<T>>(), va, nvec, transp, P0.twx(), P0.twy(), P0.twz(), s);
    VP_HIP(hipStreamSynchronize(s));
  };
  if (prec == VP_FP64) run(double{});
  else run(float{});
  VP_API_END
}

// New operator with the sources' geometry, contrast and plan whose tensor is
// sum_i w[i] fN_i, combined per precision in the stored layout (the layout is
// linear in the tensor and every source shares its prescale).
int vp_op_create_lincomb(const vp_op* ops, const double* w, int k, vp_op* out) {
  VP_API_BEGIN
  VP_REQUIRE(ops && w && out && k > 0 && k <= 64, ERR_INVALID, "bad lincomb arguments");
  const vp_op_s* o0 = ops[0];
  VP_REQUIRE(o0, ERR_INVALID, "ops[0] is NULL");
  for (int i = 1; i < k; ++i) require_combinable(o0, ops[i]);
  VP_HIP(hipSetDevice(o0->ctx->device));
  hipStream_t s = o0->ctx->stream;
  auto* op = new vp_op_s();
  try {
    op->ctx = o0->ctx;
    op->L = o0->L; op->M = o0->M; op->N = o0->N;
    op->Lp = o0->Lp; op->Mp = o0->Mp; op->Np = o0->Np;
    op->l0 = o0->l0; op->m0 = o0->m0; op->n0 = o0->n0;
    op->Lb = o0->Lb; op->Mb = o0->Mb; op->Nb = o0->Nb;
    op->V = o0->V; op->nS = o0->nS; op->structured = o0->structured; op->n = o0->n;
    op->gram = o0->gram; op->scale32 = o0->scale32; op->has_mr = o0->has_mr; op->fsym = o0->fsym;
    op->unit_contrast = o0->unit_contrast; op->all_voxels = o0->all_voxels;
    auto dup = [&](DevBuf& dst, const DevBuf& src) {
      dst.alloc(src.bytes);
      if (src.bytes) VP_HIP(hipMemcpyAsync(dst.p, src.p, src.bytes, hipMemcpyDeviceToDevice, s));
    };
    dup(op->pos, o0->pos);
    auto one = [&](auto tag) {
      using T = decltype(tag);
      const OpPrec<T>& S = prec_of<T>(const_cast<vp_op_s*>(o0));
      if (!S.present) return;
      OpPrec<T>& P = prec_of<T>(op);
      P.present = true;
      P.g = S.g;
      P.plan = S.plan;
      P.gram = S.gram;
      dup(P.mc, S.mc);
      dup(P.mr, S.mr);
      dup(P.tw, S.tw);
      P.work.alloc(S.work.bytes);
      P.work2.alloc(S.work2.bytes);
      P.xs.alloc(S.xs.bytes);
      P.ys.alloc(S.ys.bytes);
      P.fN.alloc(S.fN.bytes);
      upload_probs(op, P);
    };
    one(double{});
    one(float{});
    combine_tensors(op, ops, w, k, s);
    VP_HIP(hipStreamSynchronize(s));
  } catch (...) {
    delete op;
    throw;
  }
  *out = op;
  VP_API_END
}

// Same combination written over the tensor of an existing operator: no
// allocation, so a sweep re-interpolates one operator per wavelength.
int vp_op_lincomb_into(vp_op out, const vp_op* ops, const double* w, int k) {
  VP_API_BEGIN
  VP_REQUIRE(out && ops && w && k > 0 && k <= 64, ERR_INVALID, "bad lincomb arguments");
  for (int i = 0; i < k; ++i) {
    VP_REQUIRE(ops[i] != out, ERR_INVALID, "the target must not be one of the sources");
    require_combinable(out, ops[i]);
  }
  VP_HIP(hipSetDevice(out->ctx->device));
  combine_tensors(out, ops, w, k, out->ctx->stream);
  VP_HIP(hipStreamSynchronize(out->ctx->stream));
  VP_API_END
}

int vp_op_fp32_scale(vp_op op, double* scale) {
  VP_API_BEGIN
  VP_REQUIRE(op && scale, ERR_INVALID, "op/scale is NULL");
  *scale = op->scale32;
  VP_API_END
}

int vp_op_info(vp_op op, int64_t* out) {
  VP_API_BEGIN
  VP_REQUIRE(op && out, ERR_INVALID, "op/out is NULL");
  const int64_t v[17] = {op->L, op->M, op->N, op->l0, op->m0, op->n0, op->Lb, op->Mb, op->Nb,
                         op->p64.present ? op->p64.g.x.P : 0, op->p64.present ? op->p64.g.y.P : 0,
                         op->p64.present ? op->p64.g.z.P : 0, op->p32.present ? op->p32.g.x.P : 0,
                         op->p32.present ? op->p32.g.y.P : 0, op->p32.present ? op->p32.g.z.P : 0,
                         op->p64.present ? op->p64.plan.variant : -1, op->p32.present ? op->p32.plan.variant : -1};
  std::memcpy(out, v, sizeof(v));
  VP_API_END
}

int vp_op_size(vp_op op, int64_t* n) {
  VP_API_BEGIN
  VP_REQUIRE(op && n, ERR_INVALID, "op/n is NULL");
  *n = op->n;
  VP_API_END
}

int vp_op_destroy(vp_op op) {
  VP_API_BEGIN
  if (!op) return VP_OK;
  VP_HIP(hipSetDevice(op->ctx->device));
  (void)hipStreamSynchronize(op->ctx->stream);
  delete op;
  VP_API_END
}

}  // extern "C"
