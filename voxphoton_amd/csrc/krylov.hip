// Restarted GMRES, mixed-precision helpers and batched problem sets.
//
// Arnoldi (pgmres_iter.m:60-127) runs as fused device kernels:
//   k_dots       h1 = V^H w, ||w||^2               (FP64 partial sums per block)
//   k_axpy_dots  w -= V h1 ; ||w||^2               (mode 1)
//   k_dgks       optional second pass wanted where ||w'||^2 < eta^2 ||w||^2
//   k_dots, k_axpy_dots  h2 = V^H w, w -= V h2     (only those problems; off by
//                default: one pass per step as pgmres_iter.m:106-107)
//   k_arnoldi    H(:,k) = h1 (+ h2), ||w||, Givens rotations, residual estimate
//   k_scale      v_{k+1} = w / ||w||
// All reductions use fixed-order trees (no atomics), so a problem's result
// does not depend on which other problems share the launch.
// An augmented solve (GmresParams::rec) keeps recycled columns C ahead of the
// basis: the cycle residual is projected on C (a0 = C^H r), every
// Gram-Schmidt pass runs over [C V] (the C entries form B = C^H Aop v_k), and
// the update is V y + U (a0 - B y) with Aop U = C (GCRO).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <array>
#include <complex>
#include <map>
#include <mutex>
#include <utility>
#include <cstdio>
#include <cstdlib>
#include "eig.hpp"
#include "krylov.hpp"

namespace vp {

constexpr int KT = 256;    // threads per Krylov block
constexpr int RPT = 16;    // rows per thread
constexpr int RB = KT * RPT;
constexpr int JC = 2;      // columns per register pass

__device__ __forceinline__ double2 to_d(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ double2 to_d(double2 v) { return v; }
template <typename T> __device__ __forceinline__ C<T> from_d(double2 v);
template <> __device__ __forceinline__ float2 from_d<float>(double2 v) { return make_float2(float(v.x), float(v.y)); }
template <> __device__ __forceinline__ double2 from_d<double>(double2 v) { return v; }

template <int J>
__device__ __forceinline__ void block_reduce_store(double2 (&acc)[J], double2* out, int64_t stride, int valid) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      acc[j].x += __shfl_xor(acc[j].x, off);
      acc[j].y += __shfl_xor(acc[j].y, off);
    }
  }
  __shared__ double2 sm[KT / 64][J];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < J; ++j) sm[wave][j] = acc[j];
  __syncthreads();
  if (threadIdx.x < valid) {
    double2 s = make_double2(0.0, 0.0);
    for (int w = 0; w < KT / 64; ++w) { s.x += sm[w][threadIdx.x].x; s.y += sm[w][threadIdx.x].y; }
    out[threadIdx.x * stride] = s;
  }
  __syncthreads();
}

// Loads of the Krylov kernels are never under a branch: a row past n reads
// row n-1 (its w entry is zero, or its result is not stored) and a column
// past k reads column k-1 (its coefficient is zero, or its sum is not
// stored), so each chunk's JC x RPT loads are in flight together.
__device__ __forceinline__ int64_t krow(int64_t base, int t, int64_t n) {
  return min<int64_t>(base + t * KT + threadIdx.x, n - 1);
}

// partial dots of w (registers) with columns c0..k-1 of V (c0 a multiple of
// JC); each column's partial is the same whichever range it is computed in
template <typename T>
__device__ __forceinline__ void dots_core(const C<T>* __restrict__ Vp, int64_t vcol, int k, const double2 (&wv)[RPT],
                                          int64_t base, int64_t n, double2* part, int nblk, int c0 = 0) {
  for (int j0 = c0; j0 < k; j0 += JC) {
    C<T> v[JC][RPT];
#pragma unroll
    for (int jj = 0; jj < JC; ++jj) {
      const int64_t col = min(j0 + jj, k - 1);
#pragma unroll
      for (int t = 0; t < RPT; ++t) v[jj][t] = Vp[col * vcol + krow(base, t, n)];
    }
    double2 acc[JC];
#pragma unroll
    for (int jj = 0; jj < JC; ++jj) acc[jj] = make_double2(0.0, 0.0);
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
#pragma unroll
      for (int jj = 0; jj < JC; ++jj) {
        const double2 d = to_d(v[jj][t]);
        acc[jj].x += d.x * wv[t].x + d.y * wv[t].y;     // conj(v) * w
        acc[jj].y += d.x * wv[t].y - d.y * wv[t].x;
      }
    }
    block_reduce_store<JC>(acc, part + int64_t(j0) * nblk, nblk, min(JC, k - j0));
  }
}

// with_norm: column k of the partials also gets ||w||^2 (in .x).  Block z
// takes columns [z*cpg, (z+1)*cpg) (cpg a multiple of JC; one group of all
// columns when gridDim.z = 1), so short vectors still fill the device; the
// partials do not depend on the split.
template <typename T>
__global__ void __launch_bounds__(KT) k_dots(const C<T>* __restrict__ V, int64_t vcol, int64_t vprob, int k,
                                            const C<T>* __restrict__ w, int64_t wprob, int64_t n,
                                            double2* __restrict__ partials, int maxcols,
                                            const int* __restrict__ active, int with_norm, int cpg = 0) {
  const int p = blockIdx.y;
  if (active && !active[p]) return;
  const int c0 = cpg > 0 ? int(blockIdx.z) * cpg : 0;
  const int c1 = cpg > 0 ? min(k, c0 + cpg) : k;
  const C<T>* Vp = V + p * vprob;
  const C<T>* wp = w + p * wprob;
  const int nblk = gridDim.x;
  const int64_t base = int64_t(blockIdx.x) * RB;
  double2 wv[RPT];
#pragma unroll
  for (int t = 0; t < RPT; ++t) {
    const C<T> x = wp[krow(base, t, n)];
    wv[t] = base + t * KT + threadIdx.x < n ? to_d(x) : make_double2(0.0, 0.0);
  }
  double2* part = partials + (int64_t(p) * maxcols) * nblk + blockIdx.x;
  dots_core<T>(Vp, vcol, c1, wv, base, n, part, nblk, c0);
  if (with_norm && blockIdx.z == 0) {
    double2 acc[1] = {make_double2(0.0, 0.0)};
#pragma unroll
    for (int t = 0; t < RPT; ++t) acc[0].x += wv[t].x * wv[t].x + wv[t].y * wv[t].y;
    block_reduce_store<1>(acc, part + int64_t(k) * nblk, nblk, 1);
  }
}

// Column groups of k_dots: about 1024 workgroups in all, never more groups
// than column pairs.  Returns (groups, columns per group).
static std::pair<int, int> dots_split(int ncols, int nblk, int np) {
  const int pairs = std::max(1, (ncols + JC - 1) / JC);
  const int want = std::max(1, 1024 / std::max(1, nblk * np));
  const int g = std::min(pairs, want);
  const int cpg = ((pairs + g - 1) / g) * JC;
  return {std::max(1, (ncols + cpg - 1) / cpg), cpg};
}

// Selective reorthogonalisation (Daniel-Gragg-Kaufman-Stewart): a second
// Gram-Schmidt pass for problem p only when the first pass cancelled enough
// of w, ||w'||^2 < eta2 ||w||^2 (eta2 = 1/2 in the classical criterion;
// always = 1: every step).
__global__ void k_dgks(const double2* __restrict__ h1, const double2* __restrict__ nr, int maxcols, int k,
                       const int* __restrict__ active, int* __restrict__ reorth, int nprob, int always,
                       double eta2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob) return;
  const bool on = !active || active[p];
  const double w2 = h1[int64_t(p) * maxcols + k + 1].x;
  const double r2 = nr[int64_t(p) * maxcols].x;
  reorth[p] = (on && (always || r2 < eta2 * w2)) ? 1 : 0;
}

// w -= V h (h: FP64 per problem), then dots of the new w with V (mode 0) or
// its squared norm (mode 1).
template <typename T>
__global__ void __launch_bounds__(KT) k_axpy_dots(const C<T>* __restrict__ V, int64_t vcol, int64_t vprob, int k,
                                                 const double2* __restrict__ h, C<T>* __restrict__ w,
                                                 int64_t wprob, int64_t n, double2* __restrict__ partials,
                                                 int maxcols, int mode, const int* __restrict__ active) {
  const int p = blockIdx.y;
  if (active && !active[p]) return;
  const C<T>* Vp = V + p * vprob;
  C<T>* wp = w + p * wprob;
  const double2* hp = h + int64_t(p) * maxcols;
  const int nblk = gridDim.x;
  const int64_t base = int64_t(blockIdx.x) * RB;
  double2 wv[RPT];
#pragma unroll
  for (int t = 0; t < RPT; ++t) wv[t] = to_d(wp[krow(base, t, n)]);
  // columns in chunks of JC: the chunk's JC x RPT loads are issued before
  // any is used; each row still subtracts its terms in column order
  for (int j0 = 0; j0 < k; j0 += JC) {
    C<T> v[JC][RPT];
#pragma unroll
    for (int jj = 0; jj < JC; ++jj) {
      const int64_t col = min(j0 + jj, k - 1);
#pragma unroll
      for (int t = 0; t < RPT; ++t) v[jj][t] = Vp[col * vcol + krow(base, t, n)];
    }
#pragma unroll
    for (int jj = 0; jj < JC; ++jj) {
      if (j0 + jj < k) {
        const double2 hj = hp[j0 + jj];
#pragma unroll
        for (int t = 0; t < RPT; ++t) {
          const double2 vv = to_d(v[jj][t]);
          wv[t].x -= vv.x * hj.x - vv.y * hj.y;
          wv[t].y -= vv.x * hj.y + vv.y * hj.x;
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < RPT; ++t) {
    const int64_t i = base + t * KT + threadIdx.x;
    if (i < n) {
      const C<T> st = from_d<T>(wv[t]);
      wp[i] = st;
      wv[t] = to_d(st);
    } else {
      wv[t] = make_double2(0.0, 0.0);   // rows past n stay out of the sums below
    }
  }
  double2* part = partials + (int64_t(p) * maxcols) * nblk + blockIdx.x;
  if (mode == 0) {
    dots_core<T>(Vp, vcol, k, wv, base, n, part, nblk);
  } else {
    double2 acc[1] = {make_double2(0.0, 0.0)};
#pragma unroll
    for (int t = 0; t < RPT; ++t) acc[0].x += wv[t].x * wv[t].x + wv[t].y * wv[t].y;
    block_reduce_store<1>(acc, part, nblk, 1);
  }
}

// out[p][j] = sum over blocks of partials[p][j][blk]: one workgroup per
// (column, problem); thread t sums blocks t, t+KT, ... and a fixed tree
// combines the threads, so the order never depends on timing.
__global__ void __launch_bounds__(KT) k_reduce_cols(const double2* __restrict__ partials, int nblk, int ncol,
                                                   int maxcols, double2* __restrict__ out,
                                                   const int* __restrict__ active) {
  const int j = blockIdx.x;
  const int p = blockIdx.y;
  if (j >= ncol || (active && !active[p])) return;
  const double2* src = partials + (int64_t(p) * maxcols + j) * nblk;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  for (int b = threadIdx.x; b < nblk; b += KT) { acc[0].x += src[b].x; acc[0].y += src[b].y; }
  block_reduce_store<1>(acc, out + int64_t(p) * maxcols + j, 1, 1);
}

// Closes Arnoldi step k on the device: for each problem still stepping,
// records the step's residual estimate in hist[p][k] and the steps done,
// and clears its stepping flag once it converged or filled the basis, so
// that every later step of the cycle is a no-op for it.  The host then only
// needs the status at the end of a chunk of steps.
__global__ void k_step_end(const double* __restrict__ est, const int* __restrict__ conv, int* __restrict__ iact,
                           double* __restrict__ hist, int* __restrict__ kdone, int k, int m, int np) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np || !iact[p]) return;
  hist[int64_t(p) * m + k] = est[p];
  kdone[p] = k + 1;
  if (conv[p] || k == m - 1) iact[p] = 0;
}

// Publishes a chunk's status (the estimates of steps k0..k1-1, the steps
// done and the stepping flags) into coherent pinned host memory, then
// (after a system-scope fence) the chunk's sequence number, which the host
// polls.  The number is the device counter *dseq plus one, stored back, so a
// replayed graph of the chunk publishes a new number with the same
// arguments; the host counts chunks alike.
__global__ void k_publish(const double* __restrict__ hist, const int* __restrict__ kdone,
                          const int* __restrict__ iact, int np, int m, int k0, int k1, double* hhist, int* hkdone,
                          int* hiact, int* hseq, int* __restrict__ dseq) {
  const int w = k1 - k0;
  for (int e = threadIdx.x; e < np * w; e += blockDim.x) {
    const int p = e / w, k = k0 + e - p * w;
    reinterpret_cast<volatile double*>(hhist)[int64_t(p) * m + k] = hist[int64_t(p) * m + k];
  }
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    reinterpret_cast<volatile int*>(hkdone)[p] = kdone[p];
    reinterpret_cast<volatile int*>(hiact)[p] = iact[p];
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int seq = dseq[0] + 1;
    dseq[0] = seq;
    __threadfence_system();
    reinterpret_cast<volatile int*>(hseq)[0] = seq;
  }
}

__global__ void k_set_int(int* __restrict__ a, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) a[0] = v;
}

// Coherent pinned status buffers of the per-step publication, kept in a
// process-wide pool: a solve takes one at its start and returns it at its
// end, so repeated solves allocate no pinned memory.  The buffers live until
// the process ends.
namespace {
std::mutex g_status_mu;
std::vector<std::pair<void*, size_t>> g_status_free;

struct StatusLease {
  void* p = nullptr;
  size_t bytes = 0;
  explicit StatusLease(size_t b) {
    {
      std::lock_guard<std::mutex> lk(g_status_mu);
      for (size_t i = 0; i < g_status_free.size(); ++i)
        if (g_status_free[i].second >= b) {
          p = g_status_free[i].first;
          bytes = g_status_free[i].second;
          g_status_free.erase(g_status_free.begin() + long(i));
          return;
        }
    }
    const size_t sz = std::max<size_t>(b, 4096);
    VP_HIP(hipHostMalloc(&p, sz, hipHostMallocMapped | hipHostMallocCoherent));
    bytes = sz;
  }
  ~StatusLease() {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_status_mu);
    g_status_free.emplace_back(p, bytes);
  }
  StatusLease(const StatusLease&) = delete;
  StatusLease& operator=(const StatusLease&) = delete;
};
}  // namespace

// Workspaces of gmres_device for one shape (problems, length, restart,
// Gram-Schmidt width, recycled capacity), kept by the problem set between
// calls, and the executable graphs of Arnoldi step chunks captured over them.
// Every kernel argument of a chunk is a pointer into these buffers, the
// problem set's operator and preconditioner data, or a value fixed by the
// graph key, so a graph replays the same launches the stream path issues.
struct GmresWork {
  int64_t n = 0;
  int np = 0, m = 0, maxcols = 0, kcap = 0;
  size_t elem = 0;
  bool own_v = false;
  DevBuf Vown, t1, t2, part, H, cs, sn, g, h1, h2, nr, yv;
  DevBuf tolabs, invn, est, conv, kk, beta, bn, act, iact, ract, hist, kdone, dseq;
  DevBuf Bst, a0, coef, kcv;
  std::unique_ptr<StatusLease> pin;
  int seq = 0;                          // chunks published (host count of *dseq)
  // (k0, k1, kc, mode, second pass, recycled base column) -> graph of steps k0..k1-1 + publish
  std::map<std::array<int64_t, 6>, hipGraphExec_t> graphs;
  ~GmresWork() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
  }
};
void destroy_gmres_work(GmresWork* w) { delete w; }

// VP_GMRES_GRAPH=1 captures each chunk of Arnoldi steps into a hipGraph the
// first time the problem set runs it and replays the graph afterwards (later
// restart cycles and later inner solves of a refinement run); 0 launches the
// kernels on the stream every time.
static bool gmres_graph_enabled() {
  const char* e = std::getenv("VP_GMRES_GRAPH");
  return e && std::atoi(e) != 0;
}

// Spins until the device has published chunk ``seq``.  Past 200 ms of
// spinning it waits on the stream instead and, should the published values
// still not be visible, has them copied from the device (fallback).
template <typename F>
static void wait_published(const int* hseq, int seq, hipStream_t s, F fallback) {
  const volatile int* vs = hseq;
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned spins = 1; *vs != seq; ++spins) {
    if ((spins & 4095u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) {
      VP_HIP(hipStreamSynchronize(s));
      if (*vs != seq) fallback();
      break;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
}

// Arnoldi steps per host round trip (VP_GMRES_CHUNK, 1..64): the steps of a
// chunk are enqueued back to back and gated on the device, so a problem
// that converges inside a chunk costs at most chunk-1 steps of empty
// launches and its results are the same at every chunk length.
static int gmres_chunk() {
  const char* e = std::getenv("VP_GMRES_CHUNK");
  const int v = e ? std::atoi(e) : 4;
  return std::min(64, std::max(1, v));
}

struct ArnoldiState {
  double2* H;       // [p][(m+1) * m] column-major
  double* cs;       // [p][m]
  double2* sn;      // [p][m]
  double2* g;       // [p][m+1]
  double2* h1;      // [p][maxcols]
  double2* h2;      // [p][maxcols]
  double2* nrm;     // [p][maxcols] (entry 0: ||w||^2 in .x)
  double* tolabs;   // [p]
  double* invnorm;  // [p]
  double* est;      // [p] |g(k+1)|
  int* conv;        // [p] 1 when est < tolabs or breakdown
  int m, maxcols;
  int kc = 0;               // recycled columns ahead of the basis in h1/h2
  int kcap = 1;
  double2* Bst = nullptr;   // [p][m][kcap] C^H Aop v_k (when kc > 0)
};

__global__ void k_arnoldi(ArnoldiState st, int k, const int* __restrict__ active, int nprob,
                          const int* __restrict__ reorth) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob) return;
  if (active && !active[p]) return;
  const int m = st.m;
  double2* col = st.H + int64_t(p) * (m + 1) * m + int64_t(k) * (m + 1);
  const double2* h1 = st.h1 + int64_t(p) * st.maxcols;
  const double2* h2 = st.h2 + int64_t(p) * st.maxcols;
  const bool two = reorth[p] != 0;
  // entries 0..kc-1 are the projections on the recycled columns C
  for (int i = 0; i < st.kc; ++i)
    st.Bst[(int64_t(p) * m + k) * st.kcap + i] = two ? make_double2(h1[i].x + h2[i].x, h1[i].y + h2[i].y) : h1[i];
  h1 += st.kc;
  h2 += st.kc;
  if (two)
    for (int i = 0; i <= k; ++i) col[i] = make_double2(h1[i].x + h2[i].x, h1[i].y + h2[i].y);
  else
    for (int i = 0; i <= k; ++i) col[i] = h1[i];
  const double hn = sqrt(fmax(st.nrm[int64_t(p) * st.maxcols].x, 0.0));
  col[k + 1] = make_double2(hn, 0.0);
  double* cs = st.cs + int64_t(p) * m;
  double2* sn = st.sn + int64_t(p) * m;
  for (int i = 0; i < k; ++i) {
    const double2 a = col[i], b = col[i + 1];
    const double c = cs[i];
    const double2 s = sn[i];
    // [a; b] <- [c a + s b ; -conj(s) a + c b]
    const double2 t = make_double2(c * a.x + s.x * b.x - s.y * b.y, c * a.y + s.x * b.y + s.y * b.x);
    const double2 u = make_double2(-(s.x * a.x + s.y * a.y) + c * b.x, -(s.x * a.y - s.y * a.x) + c * b.y);
    col[i] = t;
    col[i + 1] = u;
  }
  const double2 a = col[k];
  const double bb = col[k + 1].x;   // real, >= 0
  const double aa = sqrt(a.x * a.x + a.y * a.y);
  double c;
  double2 s, r;
  if (bb == 0.0) {
    c = 1.0; s = make_double2(0.0, 0.0); r = a;
  } else if (aa == 0.0) {
    c = 0.0; s = make_double2(1.0, 0.0); r = make_double2(bb, 0.0);
  } else {
    const double rho = sqrt(aa * aa + bb * bb);
    c = aa / rho;
    s = make_double2(a.x / aa * bb / rho, a.y / aa * bb / rho);   // (a/|a|) * conj(b) / rho, b real
    r = make_double2(a.x / aa * rho, a.y / aa * rho);
  }
  cs[k] = c;
  sn[k] = s;
  col[k] = r;
  col[k + 1] = make_double2(0.0, 0.0);
  double2* g = st.g + int64_t(p) * (m + 1);
  const double2 gk = g[k];
  g[k] = make_double2(c * gk.x, c * gk.y);
  g[k + 1] = make_double2(-(s.x * gk.x + s.y * gk.y), -(s.x * gk.y - s.y * gk.x));
  const double est = sqrt(g[k + 1].x * g[k + 1].x + g[k + 1].y * g[k + 1].y);
  st.est[p] = est;
  st.invnorm[p] = hn > 0.0 ? 1.0 / hn : 0.0;
  st.conv[p] = (est < st.tolabs[p]) || !(hn > 0.0) ? 1 : 0;
}

// y = R^{-1} g for the first kk[p] columns (back substitution, FP64)
__global__ void k_backsolve(ArnoldiState st, const int* __restrict__ kk, double2* __restrict__ y, int nprob,
                            const int* __restrict__ active) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob) return;
  if (active && !active[p]) return;
  const int m = st.m;
  const int K = kk[p];
  const double2* H = st.H + int64_t(p) * (m + 1) * m;
  const double2* g = st.g + int64_t(p) * (m + 1);
  double2* yp = y + int64_t(p) * st.maxcols;
  for (int i = K - 1; i >= 0; --i) {
    double2 s = g[i];
    for (int j = i + 1; j < K; ++j) {
      const double2 h = H[int64_t(j) * (m + 1) + i];
      s.x -= h.x * yp[j].x - h.y * yp[j].y;
      s.y -= h.x * yp[j].y + h.y * yp[j].x;
    }
    const double2 d = H[int64_t(i) * (m + 1) + i];
    const double den = d.x * d.x + d.y * d.y;
    yp[i] = den > 0 ? make_double2((s.x * d.x + s.y * d.y) / den, (s.y * d.x - s.x * d.y) / den)
                    : make_double2(0.0, 0.0);
  }
}

// out (+)= sum_{j < kk[p]} V_j y_j
template <typename T>
__global__ void __launch_bounds__(KT) k_lincomb(const C<T>* __restrict__ V, int64_t vcol, int64_t vprob,
                                               const double2* __restrict__ y, const int* __restrict__ kk,
                                               int maxcols, C<T>* __restrict__ out, int64_t oprob, int64_t n,
                                               int accumulate, const int* __restrict__ active) {
  const int p = blockIdx.y;
  if (active && !active[p]) return;
  const C<T>* Vp = V + p * vprob;
  const double2* yp = y + int64_t(p) * maxcols;
  C<T>* op = out + p * oprob;
  const int K = kk[p];
  for (int64_t i = blockIdx.x * int64_t(KT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * KT) {
    double2 s = accumulate ? to_d(op[i]) : make_double2(0.0, 0.0);
    for (int j = 0; j < K; ++j) {
      const double2 v = to_d(Vp[int64_t(j) * vcol + i]);
      s.x += v.x * yp[j].x - v.y * yp[j].y;
      s.y += v.x * yp[j].y + v.y * yp[j].x;
    }
    op[i] = from_d<T>(s);
  }
}

// x[p] *= scale[p] (scale FP64, inverted when invert != 0)
template <typename T>
__global__ void k_scale(C<T>* __restrict__ x, int64_t xprob, int64_t n, const double* __restrict__ scale,
                        int invert, const int* __restrict__ active) {
  const int p = blockIdx.y;
  if (active && !active[p]) return;
  double sc = scale[p];
  if (invert) sc = sc > 0 ? 1.0 / sc : 0.0;
  C<T>* xp = x + p * xprob;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const double2 v = to_d(xp[i]);
    xp[i] = from_d<T>(make_double2(v.x * sc, v.y * sc));
  }
}

// r = b - y
template <typename T>
__global__ void k_sub(const C<T>* __restrict__ b, const C<T>* __restrict__ y, C<T>* __restrict__ r, int64_t n,
                      int nprob, const int* __restrict__ active) {
  const int64_t total = n * nprob;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
    if (active && !active[e / n]) continue;
    r[e] = csub(b[e], y[e]);
  }
}

template <typename T>
__global__ void k_copy(const C<T>* __restrict__ a, C<T>* __restrict__ b, int64_t total) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x)
    b[e] = a[e];
}

template <typename T>
__global__ void k_zero(C<T>* __restrict__ a, int64_t total) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x)
    a[e] = czero<T>();
}

__global__ void k_gmres_init(ArnoldiState st, const double* __restrict__ beta, int nprob,
                             const int* __restrict__ active) {
  const int p = blockIdx.x;
  if (p >= nprob || (active && !active[p])) return;
  const int m = st.m;
  for (int i = threadIdx.x; i <= m; i += blockDim.x)
    st.g[int64_t(p) * (m + 1) + i] = make_double2(i == 0 ? beta[p] : 0.0, 0.0);
}

// partial |x|^2 per block then per-problem total
template <typename T>
__global__ void __launch_bounds__(KT) k_norm_partial(const C<T>* __restrict__ x, int64_t xprob, int64_t n,
                                                    double2* __restrict__ partials) {
  const int p = blockIdx.y;
  const C<T>* xp = x + p * xprob;
  double2 acc[1] = {make_double2(0.0, 0.0)};
  for (int64_t i = blockIdx.x * int64_t(KT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * KT) {
    const double2 v = to_d(xp[i]);
    acc[0].x += v.x * v.x + v.y * v.y;
  }
  block_reduce_store<1>(acc, partials + int64_t(p) * gridDim.x + blockIdx.x, 1, 1);
}

__global__ void k_norm_final(const double2* __restrict__ partials, int nblk, double* __restrict__ out) {
  const int p = blockIdx.x;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += partials[int64_t(p) * nblk + b].x;
    out[p] = sqrt(s);
  }
}

template <typename T>
void dev_norms(const C<T>* x, int64_t n, int nprob, int64_t vstride, double* d_out, hipStream_t s) {
  const int nblk = int(std::min<int64_t>((n + KT - 1) / KT, 1024));
  DevBuf part(size_t(nblk) * nprob * sizeof(double2));
  hipLaunchKernelGGL(k_norm_partial<T>, dim3(nblk, nprob), dim3(KT), 0, s, x, vstride, n, part.as<double2>());
  hipLaunchKernelGGL(k_norm_final, dim3(nprob), dim3(64), 0, s, part.as<double2>(), nblk, d_out);
  VP_HIP(hipGetLastError());
  VP_HIP(hipStreamSynchronize(s));   // part is freed on return
}
template void dev_norms<float>(const float2*, int64_t, int, int64_t, double*, hipStream_t);
template void dev_norms<double>(const double2*, int64_t, int, int64_t, double*, hipStream_t);

// coef[p][i] = a0[p][i] - sum_{j < kk[p]} B[p][j][i] y[p][j], i < kc: the
// recycled part of a cycle's update (z = V y + U coef)
__global__ void k_aug_coef(const double2* __restrict__ a0, const double2* __restrict__ Bst, int kcap, int m,
                           const double2* __restrict__ y, const int* __restrict__ kk, int maxcols, int kc,
                           double2* __restrict__ coef, int nprob, const int* __restrict__ active) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob || (active && !active[p])) return;
  const int K = kk[p];
  const double2* yp = y + int64_t(p) * maxcols;
  for (int i = 0; i < kc; ++i) {
    double2 s = a0[int64_t(p) * maxcols + i];
    for (int j = 0; j < K; ++j) {
      const double2 b = Bst[(int64_t(p) * m + j) * kcap + i];
      s.x -= b.x * yp[j].x - b.y * yp[j].y;
      s.y -= b.x * yp[j].y + b.y * yp[j].x;
    }
    coef[int64_t(p) * maxcols + i] = s;
  }
}

constexpr int JO = 32;   // output columns per tall-skinny product block

// out[p][:, j] = sum_{i < kin[p]} in[p][:, i] S[p][i + ls j] for j < kout
// (S FP64 with at least ceil(kout / JO) * JO columns, zero past kout).  One
// row per thread and JO outputs per block, so each input column is read
// ceil(kout / JO) times.  The coefficients of IC inputs at a time sit in LDS
// (every lane reads the same word: a broadcast); each complex product is
// accumulated as four FP64 FMAs, the inputs in increasing order.
template <typename T>
__global__ void __launch_bounds__(KT) k_tsgemm(const C<T>* __restrict__ in, int64_t iprob,
                                              const double2* __restrict__ S, int ls, int64_t sprob,
                                              const int* __restrict__ kin, int kout, C<T>* __restrict__ out,
                                              int64_t oprob, int64_t n, const int* __restrict__ write = nullptr) {
  const int p = blockIdx.y;
  if (write && !write[p]) return;   // this problem keeps what its output columns hold
  const int j0 = blockIdx.z * JO;
  const int K = kin[p];
  const C<T>* ip = in + p * iprob;
  const double2* Sp = S + p * sprob + int64_t(j0) * ls;
  const int64_t r = int64_t(blockIdx.x) * KT + threadIdx.x;
  const int64_t rr = min<int64_t>(r, n - 1);
  double2 acc[JO];
#pragma unroll
  for (int jj = 0; jj < JO; ++jj) acc[jj] = make_double2(0.0, 0.0);
  constexpr int IC = 32, IU = 8;
  __shared__ double2 sS[JO * IC];
  for (int i0 = 0; i0 < K; i0 += IC) {
    const int ic = min(IC, K - i0);
    __syncthreads();   // the previous chunk's coefficients are no longer read
    for (int e = threadIdx.x; e < JO * IC; e += KT) {
      const int jj = e / IC, ii = e - jj * IC;
      sS[e] = ii < ic ? Sp[int64_t(jj) * ls + i0 + ii] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int iu = 0; iu < ic; iu += IU) {
      C<T> vv[IU];
#pragma unroll
      for (int u = 0; u < IU; ++u) vv[u] = ip[int64_t(i0 + min(iu + u, ic - 1)) * n + rr];
#pragma unroll
      for (int u = 0; u < IU; ++u) {
        if (iu + u < ic) {
          const double2 v = to_d(vv[u]);
          const double2* cs = sS + iu + u;
#pragma unroll
          for (int jj = 0; jj < JO; ++jj) {
            const double2 c = cs[jj * IC];
            acc[jj].x = fma(v.x, c.x, acc[jj].x);
            acc[jj].x = fma(-v.y, c.y, acc[jj].x);
            acc[jj].y = fma(v.x, c.y, acc[jj].y);
            acc[jj].y = fma(v.y, c.x, acc[jj].y);
          }
        }
      }
    }
  }
  if (r >= n) return;
  C<T>* op = out + p * oprob;
#pragma unroll
  for (int jj = 0; jj < JO; ++jj)
    if (j0 + jj < kout) op[int64_t(j0 + jj) * n + r] = from_d<T>(acc[jj]);
}

void Recycle::alloc(int64_t n, int np, int m_, size_t elem) {
  m = m_;
  kcap = m_;
  kc = 0;
  build = false;
  V.alloc(size_t(kcap + m + 1) * n * np * elem);
  U.alloc(size_t(kcap) * n * np * elem);
}

// ----- mixed precision helpers -------------------------------------------------
__global__ void k_resid_d(const double2* __restrict__ b, const double2* __restrict__ ax, double2* __restrict__ r,
                          int64_t total) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x)
    r[e] = csub(b[e], ax[e]);
}
void dev_residual_d(const double2* b, const double2* Ax, double2* r, int64_t n, int nprob, hipStream_t s) {
  const int64_t total = n * nprob;
  hipLaunchKernelGGL(k_resid_d, dim3(int(std::min<int64_t>((total + 255) / 256, 8192))), dim3(256), 0, s, b, Ax, r,
                     total);
  VP_HIP(hipGetLastError());
}
__global__ void k_scale_to_f(const double2* __restrict__ r, const double* __restrict__ nrm, float2* __restrict__ o,
                             int64_t n, int nprob, double extra) {
  const int64_t total = n * nprob;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
    const double d = nrm[e / n];
    const double sc = d > 0 ? extra / d : 0.0;
    o[e] = make_float2(float(r[e].x * sc), float(r[e].y * sc));
  }
}
void dev_scale_to_f(const double2* r, const double* d_norm, float2* out, int64_t n, int nprob, double extra,
                    hipStream_t s) {
  const int64_t total = n * nprob;
  hipLaunchKernelGGL(k_scale_to_f, dim3(int(std::min<int64_t>((total + 255) / 256, 8192))), dim3(256), 0, s, r,
                     d_norm, out, n, nprob, extra);
  VP_HIP(hipGetLastError());
}
__global__ void k_update_from_f(double2* __restrict__ x, const float2* __restrict__ d, const double* __restrict__ nrm,
                                int64_t n, int nprob, double extra, const int* __restrict__ active) {
  const int64_t total = n * nprob;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t p = e / n;
    if (active && !active[p]) continue;
    const double sc = nrm[p] * extra;
    x[e].x += double(d[e].x) * sc;
    x[e].y += double(d[e].y) * sc;
  }
}
void dev_update_from_f(double2* x, const float2* d, const double* d_norm, int64_t n, int nprob, double extra,
                       const int* active, hipStream_t s) {
  const int64_t total = n * nprob;
  hipLaunchKernelGGL(k_update_from_f, dim3(int(std::min<int64_t>((total + 255) / 256, 8192))), dim3(256), 0, s, x,
                     d, d_norm, n, nprob, extra, active);
  VP_HIP(hipGetLastError());
}

// ----- operator / preconditioner application over a problem set ------------------
bool progress_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("VP_PROGRESS");
    return e && e[0] == '1';
  }();
  return on;
}

template <typename T>
void ps_apply_A(ProblemSet<T>& ps, const C<T>* x, int64_t xs, C<T>* y, int64_t ys, const int* active) {
  VecArgs<T> va{x, y, xs, ys, active};
  launch_matvec<T>(ps.g, ps.plan, ps.probs.template as<ProbPtrs<T>>(), va, ps.nprob, 0, ps.twx, ps.twy, ps.twz,
                   ps.stream);
}
template <typename T>
void ps_apply_M(ProblemSet<T>& ps, const C<T>* x, int64_t xs, C<T>* y, int64_t ys, const int* active) {
  for (auto& part : ps.pparts) {
    PcArgs<T> a = part->pa;
    a.x = x;
    a.y = y;
    a.xstride = xs;
    a.ystride = ys;
    a.active = active;
    launch_prec<T>(part->pg, a, ps.nprob, ps.stream);
  }
}
template void ps_apply_A<float>(ProblemSet<float>&, const float2*, int64_t, float2*, int64_t, const int*);
template void ps_apply_A<double>(ProblemSet<double>&, const double2*, int64_t, double2*, int64_t, const int*);
template void ps_apply_M<float>(ProblemSet<float>&, const float2*, int64_t, float2*, int64_t, const int*);
template void ps_apply_M<double>(ProblemSet<double>&, const double2*, int64_t, double2*, int64_t, const int*);

// b[p] (+)= a[p] over problems, with separate strides
template <typename T>
__global__ void k_copy2(const C<T>* __restrict__ a, int64_t as, C<T>* __restrict__ b, int64_t bs, int64_t n,
                        int accumulate, const int* __restrict__ active) {
  const int p = blockIdx.y;
  if (active && !active[p]) return;
  const C<T>* ap = a + p * as;
  C<T>* bp = b + p * bs;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    bp[i] = accumulate ? cadd(bp[i], ap[i]) : ap[i];
}

// Width of the recycled space kept after the first inner solve of a
// refinement run: the context's setting (vp_ctx_set_recycle_width), else
// VP_IR_RECYCLE_K, else 0.  0 keeps the whole last cycle; k > 0 the k
// harmonic Ritz directions of smallest modulus (harmonic_pair).
static int recycle_k(int ctx_width) {
  if (ctx_width >= 0) return ctx_width;
  const char* e = std::getenv("VP_IR_RECYCLE_K");
  return e ? std::max(0, std::atoi(e)) : 0;
}

// ----- GMRES driver ---------------------------------------------------------------
static int grid1(int64_t total) { return int(std::min<int64_t>((total + 255) / 256, 8192)); }

template <typename T>
void gmres_device(ProblemSet<T>& ps, const C<T>* b, C<T>* x, const GmresParams& prm, GmresResult& res) {
  const int np = ps.nprob;
  const int64_t n = ps.n;
  const int m = prm.restart;
  Recycle* rec = prm.rec;
  VP_REQUIRE(!rec || rec->m == m, ERR_INVALID, "recycled space of another restart length");
  const int kcap = rec ? rec->kcap : 0;
  const int kc = rec ? rec->kc : 0;      // recycled columns C (fixed during this solve)
  const int maxcols = kcap + m + 2;
  hipStream_t s = ps.stream;
  const int nblk = int((n + RB - 1) / RB);
  const bool left = ps.has_prec && prm.side == 0;
  const bool right = ps.has_prec && prm.side == 1;

  // workspaces: the problem set's from an earlier call of the same shape, else new ones
  GmresWork* W = ps.gw.get();
  if (!W || W->n != n || W->np != np || W->m != m || W->maxcols != maxcols || W->kcap != kcap ||
      W->elem != sizeof(C<T>) || W->own_v != !rec) {
    ps.gw.reset();
    ps.gw.reset(new GmresWork);
    W = ps.gw.get();
    W->n = n;
    W->np = np;
    W->m = m;
    W->maxcols = maxcols;
    W->kcap = kcap;
    W->elem = sizeof(C<T>);
    W->own_v = !rec;
    if (!rec) W->Vown.alloc(size_t(m + 1) * n * np * sizeof(C<T>));
    W->t1.alloc(size_t(n) * np * sizeof(C<T>));
    W->t2.alloc(size_t(n) * np * sizeof(C<T>));
    W->part.alloc(size_t(np) * maxcols * nblk * sizeof(double2));
    W->H.alloc(size_t(np) * (m + 1) * m * sizeof(double2));
    W->cs.alloc(size_t(np) * m * sizeof(double));
    W->sn.alloc(size_t(np) * m * sizeof(double2));
    W->g.alloc(size_t(np) * (m + 1) * sizeof(double2));
    for (DevBuf* d : {&W->h1, &W->h2, &W->nr, &W->yv}) d->alloc(size_t(np) * maxcols * sizeof(double2));
    for (DevBuf* d : {&W->tolabs, &W->invn, &W->est, &W->beta, &W->bn}) d->alloc(np * sizeof(double));
    for (DevBuf* d : {&W->conv, &W->kk, &W->act, &W->iact, &W->ract, &W->kdone}) d->alloc(np * sizeof(int));
    W->hist.alloc(size_t(np) * m * sizeof(double));
    W->dseq.alloc(sizeof(int));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, W->dseq.as<int>(), 0);
    VP_HIP(hipGetLastError());
    // recycled part: B = C^H Aop V, a0 = C^H r per cycle, update coefficients of U
    if (kcap > 0) {
      W->Bst.alloc(size_t(np) * m * kcap * sizeof(double2));
      W->a0.alloc(size_t(np) * maxcols * sizeof(double2));
      W->coef.alloc(size_t(np) * maxcols * sizeof(double2));
      W->kcv.alloc(np * sizeof(int));
    }
    // step status: estimates hist[p][k], steps done, stepping flags, sequence number
    W->pin.reset(new StatusLease(size_t(np) * m * sizeof(double) + size_t(2 * np + 1) * sizeof(int)));
    *reinterpret_cast<volatile int*>(static_cast<int*>(W->pin->p) + (size_t(np) * m * 2 + 2 * np)) = -1;
    W->seq = 0;
  }
  DevBuf &t1 = W->t1, &t2 = W->t2, &part = W->part, &H = W->H, &cs = W->cs, &sn = W->sn, &g = W->g;
  DevBuf &h1 = W->h1, &h2 = W->h2, &nr = W->nr, &yv = W->yv;
  DevBuf &tolabs = W->tolabs, &invn = W->invn, &est = W->est, &conv = W->conv, &kk = W->kk, &beta = W->beta,
         &bn = W->bn, &act = W->act;
  DevBuf &Bst = W->Bst, &a0 = W->a0, &coef = W->coef, &kcv = W->kcv;
  if (kc > 0) {
    std::vector<int> hk(np, kc);
    VP_HIP(hipMemcpy(kcv.p, hk.data(), np * sizeof(int), hipMemcpyHostToDevice));
  }
  // Krylov storage: [C (kc of kcap columns, right-aligned) | V (m+1 columns)]
  const int64_t vprob = int64_t(kcap + m + 1) * n;
  C<T>* Vb = (rec ? rec->V.as<C<T>>() : W->Vown.as<C<T>>()) + int64_t(kcap) * n;
  C<T>* Cb = Vb - int64_t(kc) * n;   // Gram-Schmidt runs over columns Cb .. Vb + k
  ArnoldiState st{H.as<double2>(), cs.as<double>(), sn.as<double2>(), g.as<double2>(), h1.as<double2>(),
                  h2.as<double2>(), nr.as<double2>(), tolabs.as<double>(), invn.as<double>(), est.as<double>(),
                  conv.as<int>(), m, maxcols};
  st.kc = kc;
  st.kcap = std::max(kcap, 1);
  st.Bst = kc > 0 ? Bst.as<double2>() : nullptr;
  res.flag.assign(np, 1);
  res.iter_inner.assign(np, 0);
  res.iter_outer.assign(np, 0);
  res.total.assign(np, 0);
  res.relres.assign(np, 0.0);
  res.resvec.assign(np, {});
  VP_REQUIRE(prm.tol_per.empty() || int(prm.tol_per.size()) == np, ERR_INVALID, "per-problem tolerances");
  VP_REQUIRE(prm.active_in.empty() || int(prm.active_in.size()) == np, ERR_INVALID, "per-problem active flags");
  auto joins = [&](int p) { return prm.active_in.empty() || prm.active_in[size_t(p)] != 0; };
  auto tol_of = [&](int p) { return prm.tol_per.empty() ? prm.tol : prm.tol_per[size_t(p)]; };

  std::vector<int> hact(np, 1);
  for (int p = 0; p < np; ++p) hact[p] = joins(p) ? 1 : 0;
  DevBuf& ract = W->ract;                        // per-step second-pass flags
  const char* ecg = std::getenv("VP_GMRES_CGS2");  // 1: second pass at every step
  const bool cgs2_always = ecg && ecg[0] == '1';
  // One classical Gram-Schmidt pass per step, as pgmres_iter.m:106-107.
  // VP_DGKS_ETA = eta > 0 adds the selective second pass where
  // ||w'|| < eta ||w|| (eta = 1/sqrt(2) is the classical DGKS criterion).
  double eta2 = 0.0;
  if (const char* ee = std::getenv("VP_DGKS_ETA")) eta2 = std::atof(ee) * std::atof(ee);
  const bool second_pass = cgs2_always || eta2 > 0.0;
  auto upload_act = [&]() {
    VP_HIP(hipMemcpyAsync(act.p, hact.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
  };
  upload_act();
  const int* dact = act.as<int>();
  // bnorm: ||M^-1 b|| (left) or ||b||
  if (left) {
    ps_apply_M<T>(ps, b, n, t1.as<C<T>>(), n, dact);
    dev_norms<T>(t1.as<C<T>>(), n, np, n, bn.as<double>(), s);
  } else {
    dev_norms<T>(b, n, np, n, bn.as<double>(), s);
  }
  std::vector<double> hbn(np), hbeta(np);
  // per-step estimates and convergence flags: the device publishes them into
  // coherent pinned host memory, followed by the chunk's sequence number, and
  // the host spins on that number (no copy kernels, no blocking wait)
  // chunk status: step estimates hist[p][k], steps done, stepping flags
  double* hhist = static_cast<double*>(W->pin->p);
  int* hkdone = reinterpret_cast<int*>(hhist + size_t(np) * m);
  int* hiact = hkdone + np;
  int* hseq = hiact + np;
  void* dpin = nullptr;
  VP_HIP(hipHostGetDevicePointer(&dpin, W->pin->p, 0));
  double* dhhist = static_cast<double*>(dpin);
  int* dhkdone = reinterpret_cast<int*>(dhhist + size_t(np) * m);
  int* dhiact = dhkdone + np;
  int* dhseq = dhiact + np;
  DevBuf &hist = W->hist, &kdone = W->kdone;
  const int chunk = gmres_chunk();
  const bool use_graph = gmres_graph_enabled() && s != nullptr && !prof_enabled_on(s);
  VP_HIP(hipMemcpy(hbn.data(), bn.p, np * sizeof(double), hipMemcpyDeviceToHost));
  // initial residual into V_0 slot of each problem
  auto residual_into_V0 = [&](bool x_nonzero) {
    C<T>* v0 = Vb;  // problem p at Vb + p * vprob
    if (x_nonzero) {
      ps_apply_A<T>(ps, x, n, t1.as<C<T>>(), n, dact);
      hipLaunchKernelGGL(k_sub<T>, dim3(grid1(n * np)), dim3(256), 0, s, b, t1.as<C<T>>(), t2.as<C<T>>(), n, np,
                         dact);
    } else {
      hipLaunchKernelGGL(k_copy<T>, dim3(grid1(n * np)), dim3(256), 0, s, b, t2.as<C<T>>(), n * np);
    }
    VP_HIP(hipGetLastError());
    // precondition (left) or copy into the strided V_0 slots
    if (left) {
      ps_apply_M<T>(ps, t2.as<C<T>>(), n, v0, vprob, dact);
    } else {
      hipLaunchKernelGGL(k_copy2<T>, dim3(grid1(n), np), dim3(256), 0, s, t2.as<C<T>>(), n, v0, vprob, n, 0, dact);
      VP_HIP(hipGetLastError());
    }
    if (kc > 0) {
      // GCRO: a0 = C^H r, r -= C a0 (the U part of the update minimises it away)
      const auto sp = dots_split(kc, nblk, np);
      hipLaunchKernelGGL(k_dots<T>, dim3(nblk, np, sp.first), dim3(KT), 0, s, Cb, n, vprob, kc, v0, vprob, n,
                         part.as<double2>(), maxcols, dact, 0, sp.second);
      hipLaunchKernelGGL(k_reduce_cols, dim3(kc, np), dim3(KT), 0, s, part.as<double2>(), nblk, kc, maxcols,
                         a0.as<double2>(), dact);
      hipLaunchKernelGGL(k_axpy_dots<T>, dim3(nblk, np), dim3(KT), 0, s, Cb, n, vprob, kc, a0.as<double2>(), v0,
                         vprob, n, part.as<double2>(), maxcols, 1, dact);
      VP_HIP(hipGetLastError());
    }
    dev_norms<T>(v0, n, np, vprob, beta.as<double>(), s);
    VP_HIP(hipMemcpy(hbeta.data(), beta.p, np * sizeof(double), hipMemcpyDeviceToHost));
  };
  if (!prm.x0_nonzero) {
    hipLaunchKernelGGL(k_zero<T>, dim3(grid1(n * np)), dim3(256), 0, s, x, n * np);
    VP_HIP(hipGetLastError());
  }
  // x += M^-1 (V y + U coef) (right) or V y + U coef, y from the rotated
  // least squares of kk[p] columns, for the problems in mask
  auto update_x = [&](const int* mask) {
    hipLaunchKernelGGL(k_backsolve, dim3((np + 63) / 64), dim3(64), 0, s, st, kk.as<int>(), yv.as<double2>(), np,
                       mask);
    if (kc > 0)
      hipLaunchKernelGGL(k_aug_coef, dim3((np + 63) / 64), dim3(64), 0, s, a0.as<double2>(), Bst.as<double2>(), kcap,
                         m, yv.as<double2>(), kk.as<int>(), maxcols, kc, coef.as<double2>(), np, mask);
    // U right-aligned in the kcap columns, as C is ahead of the basis
    const C<T>* Ub = kc > 0 ? rec->U.template as<C<T>>() + int64_t(kcap - kc) * n : nullptr;
    const int64_t uprob = int64_t(kcap) * n;
    if (right) {
      hipLaunchKernelGGL(k_lincomb<T>, dim3(grid1(n), np), dim3(KT), 0, s, Vb, n, vprob, yv.as<double2>(),
                         kk.as<int>(), maxcols, t1.as<C<T>>(), n, n, 0, mask);
      if (kc > 0)
        hipLaunchKernelGGL(k_lincomb<T>, dim3(grid1(n), np), dim3(KT), 0, s, Ub, n, uprob, coef.as<double2>(),
                           kcv.as<int>(), maxcols, t1.as<C<T>>(), n, n, 1, mask);
      ps_apply_M<T>(ps, t1.as<C<T>>(), n, t2.as<C<T>>(), n, mask);
      hipLaunchKernelGGL(k_copy2<T>, dim3(grid1(n), np), dim3(256), 0, s, t2.as<C<T>>(), n, x, n, n, 1, mask);
    } else {
      hipLaunchKernelGGL(k_lincomb<T>, dim3(grid1(n), np), dim3(KT), 0, s, Vb, n, vprob, yv.as<double2>(),
                         kk.as<int>(), maxcols, x, n, n, 1, mask);
      if (kc > 0)
        hipLaunchKernelGGL(k_lincomb<T>, dim3(grid1(n), np), dim3(KT), 0, s, Ub, n, uprob, coef.as<double2>(),
                           kcv.as<int>(), maxcols, x, n, n, 1, mask);
    }
    VP_HIP(hipGetLastError());
  };
  // C = V Q and U = V(:, 0..K-1) R^-1 from each problem's last cycle, where
  // the Givens rotations G give G H = [R; 0] and Q = G^H (first K columns):
  // Aop V(:, 0..K-1) = V H = C R, so Aop U = C with C orthonormal
  auto build_recycled = [&](const std::vector<int>& lk) {
    // problems that build a space now (the others keep theirs)
    std::vector<int> wants(np, 1);
    if (int(rec->want.size()) == np) wants = rec->want;
    int kmax = 0;   // Arnoldi columns of the last cycle per problem, at most kcap
    for (int p = 0; p < np; ++p)
      if (wants[p]) kmax = std::max(kmax, std::min(lk[p], kcap));
    if (kmax == 0) return;
    std::vector<double2> hH(size_t(np) * (m + 1) * m), hsn(size_t(np) * m);
    std::vector<double> hcs(size_t(np) * m);
    VP_HIP(hipMemcpy(hH.data(), H.p, hH.size() * sizeof(double2), hipMemcpyDeviceToHost));
    VP_HIP(hipMemcpy(hsn.data(), sn.p, hsn.size() * sizeof(double2), hipMemcpyDeviceToHost));
    VP_HIP(hipMemcpy(hcs.data(), cs.p, hcs.size() * sizeof(double), hipMemcpyDeviceToHost));
    using cd = std::complex<double>;
    const int kdef = recycle_k(ps.recycle_width);
    // per problem: SC ((K+1) x ko) and SU (K x ko), column-major
    std::vector<std::vector<cd>> SCp(np), SUp(np);
    std::vector<int> Kp(np, 0), kop(np, 0);
    for (int p = 0; p < np; ++p) {
      const int K = wants[p] ? std::min(lk[p], kmax) : 0;
      if (K == 0) continue;
      std::vector<cd> Q(size_t(K + 1) * (K + 1), cd(0.0));     // row-major
      for (int i = 0; i <= K; ++i) Q[size_t(i) * (K + 1) + i] = 1.0;
      for (int i = K - 1; i >= 0; --i) {
        const double c = hcs[size_t(p) * m + i];
        const cd sv(hsn[size_t(p) * m + i].x, hsn[size_t(p) * m + i].y);
        for (int col = 0; col <= K; ++col) {
          const cd a = Q[size_t(i) * (K + 1) + col], b = Q[size_t(i + 1) * (K + 1) + col];
          Q[size_t(i) * (K + 1) + col] = c * a - sv * b;
          Q[size_t(i + 1) * (K + 1) + col] = std::conj(sv) * a + c * b;
        }
      }
      const double2* Hp = hH.data() + size_t(p) * (m + 1) * m;
      auto R = [&](int i, int j) { return cd(Hp[size_t(j) * (m + 1) + i].x, Hp[size_t(j) * (m + 1) + i].y); };
      bool ok = true;
      for (int i = 0; i < K; ++i) ok &= std::abs(R(i, i)) > 0.0;
      if (!ok) continue;   // singular R: this problem recycles nothing
      if (kdef > 0 && kdef < K) {
        // Hb = Q(:, 0..K-1) R, the cycle's unrotated Hessenberg matrix
        std::vector<cd> Hb(size_t(K + 1) * K, cd(0.0));
        for (int j = 0; j < K; ++j)
          for (int i = 0; i <= K; ++i) {
            cd t = 0.0;
            for (int l = 0; l <= j; ++l) t += Q[size_t(i) * (K + 1) + l] * R(l, j);
            Hb[size_t(j) * (K + 1) + i] = t;
          }
        const int kk2 = harmonic_pair(K, Hb, kdef, SCp[size_t(p)], SUp[size_t(p)]);
        if (kk2 > 0) {
          Kp[p] = K;
          kop[p] = kk2;
          continue;
        }
      }
      // the whole cycle: C = V Q(:, 0..K-1), U = V R^-1
      SCp[size_t(p)].assign(size_t(K + 1) * K, cd(0.0));
      SUp[size_t(p)].assign(size_t(K) * K, cd(0.0));
      for (int j = 0; j < K; ++j)
        for (int r = 0; r <= K; ++r) SCp[size_t(p)][size_t(j) * (K + 1) + r] = Q[size_t(r) * (K + 1) + j];
      for (int j = 0; j < K; ++j) {
        std::vector<cd> xcol(K, cd(0.0));
        for (int i = j; i >= 0; --i) {
          cd sum = (i == j) ? cd(1.0) : cd(0.0);
          for (int l = i + 1; l <= j; ++l) sum -= R(i, l) * xcol[l];
          xcol[i] = sum / R(i, i);
        }
        for (int i = 0; i <= j; ++i) SUp[size_t(p)][size_t(j) * K + i] = xcol[i];
      }
      Kp[p] = K;
      kop[p] = K;
    }
    int kn = 0;
    for (int p = 0; p < np; ++p) kn = std::max(kn, kop[p]);
    if (kn == 0) return;
    const int lsC = kmax + 1, lsU = kmax, ncol = (kn + JO - 1) / JO * JO;
    std::vector<double2> SC(size_t(np) * lsC * ncol, make_double2(0.0, 0.0)),
        SU(size_t(np) * lsU * ncol, make_double2(0.0, 0.0));
    std::vector<int> kinC(np, 0), kinU(np, 0);
    for (int p = 0; p < np; ++p) {
      const int K = Kp[p], ko = kop[p];
      if (ko == 0) continue;
      for (int j = 0; j < ko; ++j) {
        for (int r = 0; r <= K; ++r) {
          const cd v = SCp[size_t(p)][size_t(j) * (K + 1) + r];
          SC[size_t(p) * lsC * ncol + size_t(j) * lsC + r] = make_double2(v.real(), v.imag());
        }
        for (int r = 0; r < K; ++r) {
          const cd v = SUp[size_t(p)][size_t(j) * K + r];
          SU[size_t(p) * lsU * ncol + size_t(j) * lsU + r] = make_double2(v.real(), v.imag());
        }
      }
      kinC[p] = K + 1;
      kinU[p] = K;
    }
    DevBuf dSC(SC.size() * sizeof(double2)), dSU(SU.size() * sizeof(double2)), dkC(np * sizeof(int)),
        dkU(np * sizeof(int)), dwr(np * sizeof(int));
    VP_HIP(hipMemcpyAsync(dSC.p, SC.data(), dSC.bytes, hipMemcpyHostToDevice, s));
    VP_HIP(hipMemcpyAsync(dSU.p, SU.data(), dSU.bytes, hipMemcpyHostToDevice, s));
    VP_HIP(hipMemcpyAsync(dkC.p, kinC.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    VP_HIP(hipMemcpyAsync(dkU.p, kinU.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    VP_HIP(hipMemcpyAsync(dwr.p, wants.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    // C and U right-aligned in the kcap columns; problems not building keep theirs
    C<T>* Cout = Vb - int64_t(kn) * n;
    C<T>* Uout = rec->U.template as<C<T>>() + int64_t(kcap - kn) * n;
    const unsigned rblk = unsigned((n + KT - 1) / KT);
    hipLaunchKernelGGL(k_tsgemm<T>, dim3(rblk, np, ncol / JO), dim3(KT), 0, s, Vb, vprob, dSC.as<double2>(), lsC,
                       int64_t(lsC) * ncol, dkC.as<int>(), kn, Cout, vprob, n, dwr.as<int>());
    hipLaunchKernelGGL(k_tsgemm<T>, dim3(rblk, np, ncol / JO), dim3(KT), 0, s, Vb, vprob, dSU.as<double2>(), lsU,
                       int64_t(lsU) * ncol, dkU.as<int>(), kn, Uout, int64_t(kcap) * n, n, dwr.as<int>());
    VP_HIP(hipGetLastError());
    VP_HIP(hipStreamSynchronize(s));
    if (int(rec->kcp.size()) == np)
      for (int p = 0; p < np; ++p)
        if (wants[p] && kop[p] > 0) rec->kcp[p] = kn;
    rec->kc = std::max(rec->kc, kn);
    rec->build = false;
    if (int(rec->want.size()) == np) std::fill(rec->want.begin(), rec->want.end(), 0);
  };
  residual_into_V0(prm.x0_nonzero);
  std::vector<double> last(np);
  std::vector<int> proj_only(np, 0), lastk(np, 0);
  bool any_proj = false;
  for (int p = 0; p < np; ++p) {
    last[p] = relres_of(hbeta[p], hbn[p]);
    res.resvec[p].push_back(last[p]);
    hact[p] = (joins(p) && last[p] > tol_of(p)) ? 1 : 0;
    // the projection on C alone met the tolerance: x takes the U part only
    if (kc > 0 && joins(p) && !hact[p]) proj_only[p] = 1;
    any_proj |= proj_only[p] != 0;
  }
  if (any_proj) {
    std::vector<int> zero(np, 0);
    DevBuf pm(np * sizeof(int));
    VP_HIP(hipMemcpyAsync(kk.p, zero.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    VP_HIP(hipMemcpyAsync(pm.p, proj_only.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    update_x(pm.as<int>());
    VP_HIP(hipStreamSynchronize(s));
  }
  {
    std::vector<double> ta(np);
    for (int p = 0; p < np; ++p) ta[p] = tol_of(p) * hbn[p];
    VP_HIP(hipMemcpy(tolabs.p, ta.data(), np * sizeof(double), hipMemcpyHostToDevice));
  }
  int outit = 0;
  std::vector<int> hk(np, 0);
  while (outit < prm.maxit) {
    bool any = false;
    for (int p = 0; p < np; ++p) any |= hact[p] != 0;
    if (!any) break;
    ++outit;
    upload_act();
    std::vector<int> inner(hact);   // active within this cycle
    for (int p = 0; p < np; ++p) {
      if (hact[p]) res.iter_outer[p] = outit;
      hk[p] = 0;
    }
    VP_HIP(hipMemcpyAsync(W->iact.p, inner.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    int* dia = W->iact.as<int>();
    hipLaunchKernelGGL(k_scale<T>, dim3(grid1(n), np), dim3(256), 0, s, Vb, vprob, n, beta.as<double>(), 1, dia);
    hipLaunchKernelGGL(k_gmres_init, dim3(np), dim3(64), 0, s, st, beta.as<double>(), np, dia);
    VP_HIP(hipGetLastError());
    // Arnoldi steps in chunks: each step's kernels are gated on the device
    // stepping flags dia, which k_step_end clears as problems stop
    auto step = [&](int k) {
      C<T>* vk = Vb + int64_t(k) * n;
      C<T>* w = Vb + int64_t(k + 1) * n;
      if (right) {          // w = A M^-1 v_k
        ps_apply_M<T>(ps, vk, vprob, t1.as<C<T>>(), n, dia);
        ps_apply_A<T>(ps, t1.as<C<T>>(), n, w, vprob, dia);
      } else if (left) {    // w = M^-1 A v_k
        ps_apply_A<T>(ps, vk, vprob, t1.as<C<T>>(), n, dia);
        ps_apply_M<T>(ps, t1.as<C<T>>(), n, w, vprob, dia);
      } else {
        ps_apply_A<T>(ps, vk, vprob, w, vprob, dia);
      }
      // first classical Gram-Schmidt pass over [C, V]: h1 = [C V]^H w (and ||w||^2),
      // w' = w - [C V] h1, ||w'||^2
      const int nc = kc + k + 1;
      const auto sp = dots_split(nc, nblk, np);
      hipLaunchKernelGGL(k_dots<T>, dim3(nblk, np, sp.first), dim3(KT), 0, s, Cb, n, vprob, nc, w, vprob, n,
                         part.as<double2>(), maxcols, dia, 1, sp.second);
      hipLaunchKernelGGL(k_reduce_cols, dim3(nc + 1, np), dim3(KT), 0, s, part.as<double2>(), nblk, nc + 1, maxcols,
                         h1.as<double2>(), dia);
      hipLaunchKernelGGL(k_axpy_dots<T>, dim3(nblk, np), dim3(KT), 0, s, Cb, n, vprob, nc, h1.as<double2>(), w,
                         vprob, n, part.as<double2>(), maxcols, 1, dia);
      hipLaunchKernelGGL(k_reduce_cols, dim3(1, np), dim3(KT), 0, s, part.as<double2>(), nblk, 1, maxcols,
                         nr.as<double2>(), dia);
      // optional second pass only where the first cancelled (blocks of other problems exit at once)
      hipLaunchKernelGGL(k_dgks, dim3((np + 63) / 64), dim3(64), 0, s, h1.as<double2>(), nr.as<double2>(), maxcols,
                         nc - 1, dia, ract.as<int>(), np, cgs2_always ? 1 : 0, eta2);
      if (second_pass) {
        hipLaunchKernelGGL(k_dots<T>, dim3(nblk, np, sp.first), dim3(KT), 0, s, Cb, n, vprob, nc, w, vprob, n,
                           part.as<double2>(), maxcols, ract.as<int>(), 0, sp.second);
        hipLaunchKernelGGL(k_reduce_cols, dim3(nc, np), dim3(KT), 0, s, part.as<double2>(), nblk, nc,
                           maxcols, h2.as<double2>(), ract.as<int>());
        hipLaunchKernelGGL(k_axpy_dots<T>, dim3(nblk, np), dim3(KT), 0, s, Cb, n, vprob, nc, h2.as<double2>(),
                           w, vprob, n, part.as<double2>(), maxcols, 1, ract.as<int>());
        hipLaunchKernelGGL(k_reduce_cols, dim3(1, np), dim3(KT), 0, s, part.as<double2>(), nblk, 1, maxcols,
                           nr.as<double2>(), ract.as<int>());
      }
      hipLaunchKernelGGL(k_arnoldi, dim3((np + 63) / 64), dim3(64), 0, s, st, k, dia, np, ract.as<int>());
      hipLaunchKernelGGL(k_scale<T>, dim3(grid1(n), np), dim3(256), 0, s, w, vprob, n, invn.as<double>(), 0, dia);
      VP_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_step_end, dim3((np + 63) / 64), dim3(64), 0, s, est.as<double>(), conv.as<int>(), dia,
                         hist.as<double>(), kdone.as<int>(), k, m, np);
      VP_HIP(hipGetLastError());
    };
    // steps k0..k1-1 and the chunk's publication, on the stream or as the
    // problem set's captured graph of that chunk
    auto issue_chunk = [&](int k0, int k1) {
      for (int k = k0; k < k1; ++k) step(k);
      hipLaunchKernelGGL(k_publish, dim3(1), dim3(256), 0, s, hist.as<double>(), kdone.as<int>(), dia, np, m, k0, k1,
                         dhhist, dhkdone, dhiact, dhseq, W->dseq.as<int>());
      VP_HIP(hipGetLastError());
    };
    auto run_chunk = [&](int k0, int k1) {
      if (!use_graph) {
        issue_chunk(k0, k1);
        return;
      }
      const std::array<int64_t, 6> key{k0, k1, kc, right ? 2 : (left ? 1 : 0), second_pass ? 1 : 0,
                                       static_cast<int64_t>(reinterpret_cast<intptr_t>(Vb))};
      auto it = W->graphs.find(key);
      if (it == W->graphs.end()) {
        VP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipGraph_t gr = nullptr;
        try {
          issue_chunk(k0, k1);
        } catch (...) {
          (void)hipStreamEndCapture(s, &gr);
          if (gr) (void)hipGraphDestroy(gr);
          throw;
        }
        VP_HIP(hipStreamEndCapture(s, &gr));
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        VP_HIP(e);
        it = W->graphs.emplace(key, ex).first;
      }
      VP_HIP(hipGraphLaunch(it->second, s));
    };
    for (int k0 = 0; k0 < m;) {
      const int k1 = std::min(m, k0 + chunk);
      run_chunk(k0, k1);
      const int seq = ++W->seq;
      wait_published(hseq, seq, s, [&] {
        VP_HIP(hipMemcpy(hhist, hist.p, size_t(np) * m * sizeof(double), hipMemcpyDeviceToHost));
        VP_HIP(hipMemcpy(hkdone, kdone.p, np * sizeof(int), hipMemcpyDeviceToHost));
        VP_HIP(hipMemcpy(hiact, dia, np * sizeof(int), hipMemcpyDeviceToHost));
      });
      bool still = false;
      for (int p = 0; p < np; ++p) {
        if (!inner[p]) continue;
        const int kd = hkdone[p];
        for (int j = k0; j < kd; ++j) {
          res.total[p] += 1;
          last[p] = relres_of(hhist[size_t(p) * m + j], hbn[p]);
          res.resvec[p].push_back(last[p]);
        }
        hk[p] = kd;
        lastk[p] = kd;
        res.iter_inner[p] = kd;
        if (hiact[p]) still = true;
        else inner[p] = 0;
      }
      if (!still) break;
      k0 = k1;
    }
    // solve the least squares and update x for every problem active in this cycle
    VP_HIP(hipMemcpyAsync(kk.p, hk.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    update_x(dact);
    VP_HIP(hipStreamSynchronize(s));   // the update is done before the host reuses the cycle's buffers
    if (progress_enabled()) {
      double worst = 0.0;
      int nact = 0;
      for (int p = 0; p < np; ++p)
        if (hact[p]) { worst = std::max(worst, last[p]); ++nact; }
      std::fprintf(stderr, "[vp] gmres cycle %d: %d problem(s), worst estimated relres %.3e\n", outit, nact, worst);
      std::fflush(stderr);
    }
    // restart: fresh residual (pgmres_iter.m:150 carries it by recurrence)
    for (int p = 0; p < np; ++p) hact[p] = (hact[p] && last[p] > tol_of(p)) ? 1 : 0;
    bool more = false;
    for (int p = 0; p < np; ++p) more |= hact[p] != 0;
    if (!more || outit >= prm.maxit) break;
    upload_act();
    residual_into_V0(true);
  }
  if (rec && rec->build && (kc == 0 || int(rec->want.size()) == np)) build_recycled(lastk);
  if (!prm.final_residual) {
    // caller recomputes the residual itself: report the recurrence estimate
    for (int p = 0; p < np; ++p) {
      res.relres[p] = last[p];
      res.flag[p] = last[p] <= tol_of(p) ? 0 : 1;
    }
    return;
  }
  // final unpreconditioned relative residual in FP64 from the T data
  std::vector<int> all(np, 1);
  VP_HIP(hipMemcpyAsync(act.p, all.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
  ps_apply_A<T>(ps, x, n, t1.as<C<T>>(), n, act.as<int>());
  hipLaunchKernelGGL(k_sub<T>, dim3(grid1(n * np)), dim3(256), 0, s, b, t1.as<C<T>>(), t2.as<C<T>>(), n, np,
                     nullptr);
  VP_HIP(hipGetLastError());
  dev_norms<T>(t2.as<C<T>>(), n, np, n, beta.as<double>(), s);
  dev_norms<T>(b, n, np, n, bn.as<double>(), s);
  VP_HIP(hipMemcpy(hbeta.data(), beta.p, np * sizeof(double), hipMemcpyDeviceToHost));
  VP_HIP(hipMemcpy(hbn.data(), bn.p, np * sizeof(double), hipMemcpyDeviceToHost));
  for (int p = 0; p < np; ++p) {
    res.relres[p] = relres_of(hbeta[p], hbn[p]);
    res.flag[p] = res.relres[p] <= tol_of(p) ? 0 : 1;
  }
}
template void gmres_device<float>(ProblemSet<float>&, const float2*, float2*, const GmresParams&, GmresResult&);
template void gmres_device<double>(ProblemSet<double>&, const double2*, double2*, const GmresParams&, GmresResult&);

// ----- flexible outer refinement ---------------------------------------------
// o[p] = v[p] rounded to complex64
__global__ void k_cvt_d2f(const double2* __restrict__ v, int64_t vprob, float2* __restrict__ o, int64_t oprob,
                          int64_t n, const int* __restrict__ active) {
  const int p = blockIdx.y;
  if (active && !active[p]) return;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const double2 a = v[p * vprob + i];
    o[p * oprob + i] = make_float2(float(a.x), float(a.y));
  }
}

// o[p] = scale * z[p] in complex128
__global__ void k_cvt_f2d(const float2* __restrict__ z, int64_t zprob, double2* __restrict__ o, int64_t oprob,
                          int64_t n, double scale, const int* __restrict__ active) {
  const int p = blockIdx.y;
  if (active && !active[p]) return;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float2 a = z[p * zprob + i];
    o[p * oprob + i] = make_double2(scale * double(a.x), scale * double(a.y));
  }
}

// x[p] += scale * sum_{j < kk[p]} y[p][j] Z[p][j]: complex64 columns, FP64
// coefficients and accumulation, columns in order for every row
__global__ void __launch_bounds__(KT) k_lincomb_zx(const float2* __restrict__ Z, int64_t zcol, int64_t zprob,
                                                  const double2* __restrict__ y, const int* __restrict__ kk,
                                                  int maxcols, double scale, double2* __restrict__ x, int64_t n,
                                                  const int* __restrict__ active) {
  const int p = blockIdx.y;
  if (active && !active[p]) return;
  const float2* Zp = Z + p * zprob;
  const double2* yp = y + int64_t(p) * maxcols;
  double2* xp = x + p * n;
  const int K = kk[p];
  for (int64_t i = blockIdx.x * int64_t(KT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * KT) {
    double2 s = make_double2(0.0, 0.0);
    for (int j = 0; j < K; ++j) {
      const float2 zf = Zp[int64_t(j) * zcol + i];
      s.x += double(zf.x) * yp[j].x - double(zf.y) * yp[j].y;
      s.y += double(zf.x) * yp[j].y + double(zf.y) * yp[j].x;
    }
    xp[i] = make_double2(xp[i].x + scale * s.x, xp[i].y + scale * s.y);
  }
}

// r[p] = b[p] - ax[p] into column 0 of the outer basis (stride vprob)
__global__ void k_resid_col(const double2* __restrict__ b, const double2* __restrict__ ax, double2* __restrict__ r,
                            int64_t vprob, int64_t n, const int* __restrict__ active) {
  const int p = blockIdx.y;
  if (active && !active[p]) return;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const double2 bb = b[p * n + i], a = ax[p * n + i];
    r[p * vprob + i] = make_double2(bb.x - a.x, bb.y - a.y);
  }
}

void fgmres_ir_set(ProblemSet<double>& p64, ProblemSet<float>& p32, const double2* b, double2* x, double scale32,
                   double tol, int restart, double inner_tol, int inner_maxit, int max_outer, int outer_m, int side,
                   std::vector<IrProblem>& st, Recycle* rec) {
  using cd = std::complex<double>;
  const int np = p64.nprob;
  const int64_t n = p64.n;
  hipStream_t s = p64.stream;
  const int mo = std::max(1, std::min(outer_m, std::max(1, max_outer)));
  const int maxc = mo + 2;
  const int nblk = int((n + RB - 1) / RB);
  const int64_t vprob = int64_t(mo + 1) * n, zprob = int64_t(mo) * n;
  DevBuf Vo(size_t(vprob) * np * sizeof(double2)), Zo(size_t(zprob) * np * sizeof(float2));
  DevBuf r32(size_t(n) * np * sizeof(float2)), d32(size_t(n) * np * sizeof(float2)),
      t64(size_t(n) * np * sizeof(double2));
  DevBuf part(size_t(np) * maxc * nblk * sizeof(double2)), h1(size_t(np) * maxc * sizeof(double2)),
      h2(size_t(np) * maxc * sizeof(double2)), nr(size_t(np) * maxc * sizeof(double2)),
      yv(size_t(np) * maxc * sizeof(double2));
  DevBuf kk(np * sizeof(int)), act(np * sizeof(int)), cact(np * sizeof(int)), beta(np * sizeof(double)),
      bnd(np * sizeof(double)), invn(np * sizeof(double));
  double2* V = Vo.as<double2>();
  float2* Z = Zo.as<float2>();
  const dim3 gv(unsigned(grid1(n)), unsigned(np));
  dev_norms<double>(b, n, np, n, bnd.as<double>(), s);
  std::vector<double> hbn(np), hbeta(np), hinv(np);
  VP_HIP(hipMemcpy(hbn.data(), bnd.p, np * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<int> hact(np);
  for (int p = 0; p < np; ++p) hact[p] = st[p].done ? 0 : 1;
  // host least squares per problem: H column-major (mo+1) x mo, Givens, g
  std::vector<cd> H(size_t(np) * (mo + 1) * mo), sn(size_t(np) * mo), g(size_t(np) * (mo + 1));
  std::vector<double> cs(size_t(np) * mo);
  std::vector<double2> hh1(size_t(np) * maxc), hh2(size_t(np) * maxc), hnr(size_t(np) * maxc),
      hy(size_t(np) * maxc);
  bool first = true;
  for (;;) {
    // true residual of the current iterate: r = b - A x (before any step x = 0 and r = b)
    VP_HIP(hipMemcpyAsync(act.p, hact.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    bool all_first = first;
    for (int p = 0; p < np; ++p) all_first &= !hact[p] || st[p].steps == 0;
    if (all_first) {
      hipLaunchKernelGGL(k_copy2<double>, gv, dim3(256), 0, s, b, n, V, vprob, n, 0, act.as<int>());
    } else {
      ps_apply_A<double>(p64, x, n, t64.as<double2>(), n, act.as<int>());
      hipLaunchKernelGGL(k_resid_col, gv, dim3(256), 0, s, b, t64.as<double2>(), V, vprob, n, act.as<int>());
    }
    VP_HIP(hipGetLastError());
    dev_norms<double>(V, n, np, vprob, beta.as<double>(), s);
    VP_HIP(hipMemcpy(hbeta.data(), beta.p, np * sizeof(double), hipMemcpyDeviceToHost));
    bool any = false;
    for (int p = 0; p < np; ++p) {
      if (!hact[p]) continue;
      IrProblem& q = st[p];
      q.relres = relres_of(hbeta[p], hbn[p]);
      // inside a cycle the history holds the FGMRES estimate of each step; the
      // true residual of the same iterate replaces it at the cycle's end
      if (first || q.hist.empty()) q.hist.push_back(q.relres);
      else q.hist.back() = q.relres;
      if (q.relres <= tol || !std::isfinite(q.relres)) {
        q.done = true;
        q.flag = q.relres <= tol ? 0 : 1;
        hact[p] = 0;
      } else if (q.steps >= max_outer) {
        q.done = true;
        q.flag = 1;
        hact[p] = 0;
      } else {
        any = true;
      }
    }
    first = false;
    if (progress_enabled()) {
      for (int p = 0; p < np; ++p)
        std::fprintf(stderr, "[vp] flexible refinement: problem %d step %d, FP64 relres %.3e, inner iterations %d%s\n",
                     p, st[p].steps, st[p].relres, st[p].inner, st[p].done ? " (done)" : "");
      std::fflush(stderr);
    }
    if (!any) break;
    // v_0 = r / ||r||
    VP_HIP(hipMemcpyAsync(act.p, hact.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scale<double>, gv, dim3(256), 0, s, V, vprob, n, beta.as<double>(), 1, act.as<int>());
    VP_HIP(hipGetLastError());
    std::fill(H.begin(), H.end(), cd(0.0));
    std::fill(g.begin(), g.end(), cd(0.0));
    for (int p = 0; p < np; ++p) g[size_t(p) * (mo + 1)] = hbeta[p];
    std::vector<int> cyc(hact), kdone(np, 0);   // in this cycle; steps taken
    std::vector<int> step_act(hact);            // still stepping
    for (int j = 0; j < mo; ++j) {
      bool stepping = false;
      for (int p = 0; p < np; ++p) stepping |= step_act[p] != 0;
      if (!stepping) break;
      VP_HIP(hipMemcpyAsync(act.p, step_act.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
      const int* da = act.as<int>();
      // z_j = inner FP32 solve of (A/s) d = v_j
      hipLaunchKernelGGL(k_cvt_d2f, gv, dim3(256), 0, s, V + int64_t(j) * n, vprob, r32.as<float2>(), n, n, da);
      VP_HIP(hipGetLastError());
      GmresParams prm;
      prm.restart = restart;
      prm.tol = inner_tol;
      prm.maxit = inner_maxit;
      prm.side = side;
      prm.final_residual = false;
      prm.active_in = step_act;
      prm.tol_per.assign(np, inner_tol);
      for (int p = 0; p < np; ++p)
        if (step_act[p]) {
          // the outer residual estimate before this step decides how deep the inner solve goes
          const double cur = std::abs(g[size_t(p) * (mo + 1) + j]) / (hbn[p] > 0 ? hbn[p] : 1.0);
          if (cur > 0) {
            const double need = std::min(0.5, 0.5 * tol / cur);
            prm.tol_per[p] = need >= inner_tol / 8 ? need : inner_tol;
          }
        }
      if (rec) {
        prm.rec = rec;
        if (rec->want.empty() ? rec->kc == 0 : std::any_of(rec->want.begin(), rec->want.end(), [](int v) { return v != 0; }))
          rec->build = true;
      }
      GmresResult gr;
      gmres_device<float>(p32, r32.as<float2>(), d32.as<float2>(), prm, gr);
      for (int p = 0; p < np; ++p)
        if (step_act[p]) {
          st[p].inner += gr.total[p];
          st[p].steps += 1;
        }
      // Z_j = d (complex64), w = A (d / s) into v_{j+1}
      double2* w = V + int64_t(j + 1) * n;
      hipLaunchKernelGGL(k_copy2<float>, gv, dim3(256), 0, s, d32.as<float2>(), n, Z + int64_t(j) * n, zprob, n, 0,
                         da);
      hipLaunchKernelGGL(k_cvt_f2d, gv, dim3(256), 0, s, d32.as<float2>(), n, t64.as<double2>(), n, n, 1.0 / scale32,
                         da);
      VP_HIP(hipGetLastError());
      ps_apply_A<double>(p64, t64.as<double2>(), n, w, vprob, da);
      // classical Gram-Schmidt, two passes, against v_0..v_j
      const int nc = j + 1;
      const auto sp = dots_split(nc, nblk, np);
      hipLaunchKernelGGL(k_dots<double>, dim3(nblk, np, sp.first), dim3(KT), 0, s, V, n, vprob, nc, w, vprob, n,
                         part.as<double2>(), maxc, da, 0, sp.second);
      hipLaunchKernelGGL(k_reduce_cols, dim3(nc, np), dim3(KT), 0, s, part.as<double2>(), nblk, nc, maxc,
                         h1.as<double2>(), da);
      hipLaunchKernelGGL(k_axpy_dots<double>, dim3(nblk, np), dim3(KT), 0, s, V, n, vprob, nc, h1.as<double2>(), w,
                         vprob, n, part.as<double2>(), maxc, 0, da);
      hipLaunchKernelGGL(k_reduce_cols, dim3(nc, np), dim3(KT), 0, s, part.as<double2>(), nblk, nc, maxc,
                         h2.as<double2>(), da);
      hipLaunchKernelGGL(k_axpy_dots<double>, dim3(nblk, np), dim3(KT), 0, s, V, n, vprob, nc, h2.as<double2>(), w,
                         vprob, n, part.as<double2>(), maxc, 1, da);
      hipLaunchKernelGGL(k_reduce_cols, dim3(1, np), dim3(KT), 0, s, part.as<double2>(), nblk, 1, maxc,
                         nr.as<double2>(), da);
      VP_HIP(hipGetLastError());
      VP_HIP(hipMemcpyAsync(hh1.data(), h1.p, hh1.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
      VP_HIP(hipMemcpyAsync(hh2.data(), h2.p, hh2.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
      VP_HIP(hipMemcpyAsync(hnr.data(), nr.p, hnr.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
      VP_HIP(hipStreamSynchronize(s));
      for (int p = 0; p < np; ++p) {
        hinv[p] = 0.0;
        if (!step_act[p]) continue;
        cd* Hc = H.data() + (size_t(p) * mo + j) * (mo + 1);   // column j of problem p
        for (int i = 0; i <= j; ++i) {
          const double2 a = hh1[size_t(p) * maxc + i], c = hh2[size_t(p) * maxc + i];
          Hc[i] = cd(a.x + c.x, a.y + c.y);
        }
        const double hn = std::sqrt(std::max(0.0, hnr[size_t(p) * maxc].x));
        Hc[j + 1] = hn;
        // earlier rotations, then the one that zeroes H(j+1, j)
        for (int i = 0; i < j; ++i) {
          const double c = cs[size_t(p) * mo + i];
          const cd sv = sn[size_t(p) * mo + i];
          const cd a = Hc[i], bb = Hc[i + 1];
          Hc[i] = c * a + sv * bb;
          Hc[i + 1] = -std::conj(sv) * a + c * bb;
        }
        const cd a = Hc[j];
        const double den = std::sqrt(std::norm(a) + hn * hn);
        double c = 1.0;
        cd sv(0.0);
        if (den > 0) {
          const double aa = std::abs(a);
          if (aa > 0) {
            c = aa / den;
            sv = (a / aa) * hn / den;
          } else {
            c = 0.0;
            sv = 1.0;
          }
        }
        // [c s; -conj(s) c] maps (a, hn) to (den * a/|a|, 0)
        cs[size_t(p) * mo + j] = c;
        sn[size_t(p) * mo + j] = sv;
        Hc[j] = c * a + sv * cd(hn);
        Hc[j + 1] = 0.0;
        cd* gp = g.data() + size_t(p) * (mo + 1);
        const cd gj = gp[j];
        gp[j] = c * gj;
        gp[j + 1] = -std::conj(sv) * gj;
        kdone[p] = j + 1;
        const double est = relres_of(std::abs(gp[j + 1]), hbn[p]);
        st[p].hist.push_back(est);
        if (!(est > tol) || hn == 0.0 || st[p].steps >= max_outer || !std::isfinite(est)) step_act[p] = 0;
        hinv[p] = hn > 0 ? 1.0 / hn : 0.0;
      }
      // v_{j+1} = w / ||w||
      VP_HIP(hipMemcpyAsync(invn.p, hinv.data(), np * sizeof(double), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_scale<double>, gv, dim3(256), 0, s, w, vprob, n, invn.as<double>(), 0, da);
      VP_HIP(hipGetLastError());
    }
    // x += Z y / s, y = R^-1 g (back substitution over each problem's steps)
    for (int p = 0; p < np; ++p) {
      const int K = kdone[p];
      const cd* Hp = H.data() + size_t(p) * mo * (mo + 1);
      const cd* gp = g.data() + size_t(p) * (mo + 1);
      std::vector<cd> yy(std::max(K, 1), cd(0.0));
      for (int i = K - 1; i >= 0; --i) {
        cd sum = gp[i];
        for (int l = i + 1; l < K; ++l) sum -= Hp[size_t(l) * (mo + 1) + i] * yy[l];
        const cd d = Hp[size_t(i) * (mo + 1) + i];
        yy[i] = std::abs(d) > 0 ? sum / d : cd(0.0);
      }
      for (int i = 0; i < maxc; ++i)
        hy[size_t(p) * maxc + i] = i < K ? make_double2(yy[i].real(), yy[i].imag()) : make_double2(0.0, 0.0);
    }
    VP_HIP(hipMemcpyAsync(yv.p, hy.data(), hy.size() * sizeof(double2), hipMemcpyHostToDevice, s));
    VP_HIP(hipMemcpyAsync(kk.p, kdone.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    VP_HIP(hipMemcpyAsync(cact.p, cyc.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_lincomb_zx, gv, dim3(KT), 0, s, Z, n, zprob, yv.as<double2>(), kk.as<int>(), maxc,
                       1.0 / scale32, x, n, cact.as<int>());
    VP_HIP(hipGetLastError());
    VP_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < np; ++p) hact[p] = cyc[p];
  }
}

}  // namespace vp
