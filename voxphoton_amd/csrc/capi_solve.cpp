// C ABI: preconditioners, GMRES, mixed-precision iterative refinement and
// batched solves.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>
#include "krylov.hpp"
#include "capi_common.hpp"

using namespace vp;


namespace vp {

// ----------------------------------------------------------------------------
// Preconditioner construction
// ----------------------------------------------------------------------------
__global__ void k_scale_c128(double2* a, int64_t n, double s);

static void prec_geometry(vp_prec_s* p, int level) {
  PcGeom& g = p->g;
  g.level = level;
  const int dev = p->ctx->device;
  const int lds_max = max_lds_bytes(dev);
  VP_REQUIRE(factorize(g.L, true, g.px, 1), ERR_UNSUPPORTED,
             "preconditioner box length L has a prime factor above the supported radices");
  if (level == 2)
    VP_REQUIRE(factorize(g.M, true, g.py, 1), ERR_UNSUPPORTED,
               "preconditioner box length M has a prime factor above the supported radices");
  const size_t s = 16;   // size for the FP64 path bounds both precisions
  g.B = 1;
  for (int B : {16, 8, 4, 2, 1}) {
    if (B > 1 && B > g.M) continue;
    const int ld = g.L + (B > 1 ? 1 : 0);
    if (2 * size_t(B) * ld * s <= size_t(lds_max)) {
      g.B = B;
      g.ld = ld;
      break;
    }
  }
  g.ld = g.L + (g.B > 1 ? 1 : 0);
  VP_REQUIRE(2 * size_t(g.B) * g.ld * s <= size_t(lds_max), ERR_UNSUPPORTED, "preconditioner x pencil exceeds LDS");
  g.lds1 = 2 * size_t(g.B) * g.ld * s;
  g.tx = 256;
  g.t2 = 512;
  g.lds2 = size_t(2) * 3 * g.N * g.M * s;
  g.psplit = 0;
  g.pquad = 0;
  g.pct = 0;
  g.Brow = 0;
  if (level == 2) {
    // y-row DFT passes (split and quadrant paths): Brow rows per workgroup
    for (int B : {16, 8, 4, 2, 1}) {
      const int ld = g.M + (B > 1 ? 1 : 0);
      if (2 * size_t(B) * ld * s <= size_t(std::min(lds_max, 40 * 1024)) || B == 1) {
        g.Brow = B;
        g.ldrow = ld;
        g.lds_row = 2 * size_t(B) * ld * s;
        break;
      }
    }
    if (g.lds_row > size_t(lds_max)) g.Brow = 0;
  }
  const char* e = std::getenv("VP_PC_SPLIT");
  if (level == 2 && (g.lds2 > size_t(lds_max) || (e && e[0] == '1'))) {
    g.psplit = 1;
    VP_REQUIRE(g.Brow > 0, ERR_UNSUPPORTED, "preconditioner y rows exceed LDS");
    g.TK = 1;
    for (int tk : {64, 32, 16, 8, 4, 2, 1})
      if ((tk <= g.M || tk == 1) && size_t(3) * g.N * tk * s <= 48 * 1024) {
        g.TK = tk;
        break;
      }
    g.lds_tile = size_t(3) * g.N * g.TK * s;
    VP_REQUIRE(g.lds_tile <= size_t(lds_max), ERR_UNSUPPORTED, "preconditioner block column exceeds LDS");
  }
}

static vp_prec_s* prec_begin(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const int64_t* idx, int64_t n,
                             int level) {
  VP_REQUIRE(ctx, ERR_INVALID, "ctx is NULL");
  VP_REQUIRE(L > 0 && M > 0 && N > 0, ERR_INVALID, "box extents must be positive");
  VP_HIP(hipSetDevice(ctx->device));
  auto* p = new vp_prec_s();
  try {
    p->ctx = ctx;
    p->g.L = int(L);
    p->g.M = int(M);
    p->g.N = int(N);
    p->g.V = L * M * N;
    p->g.bsym = 0;
    p->g.Hx = int(L);
    p->g.Hy = int(M);
    std::vector<int> pos;
    index_map(p->g.V, idx, n, pos, p->g.structured, p->g.nS);
    p->nbox = n;
    p->g.n_outer = 0;
    p->pos.alloc(pos.size() * sizeof(int));
    VP_HIP(hipMemcpy(p->pos.p, pos.data(), pos.size() * sizeof(int), hipMemcpyHostToDevice));
    p->pos_hash = map_hash(pos);
    prec_geometry(p, level);
  } catch (...) {
    delete p;
    throw;
  }
  return p;
}

template <typename T>
static void prec_alloc_prec(vp_prec_s* p, PrecPrec<T>& P) {
  P.present = true;
  const PcGeom& g = p->g;
  std::vector<C<T>> tw;
  for (int dimP : {g.L, g.M}) {
    for (int k = 0; k < dimP; ++k) {
      SinCos v = sincos_turn(k, dimP);
      tw.push_back(mk<T>(T(v.c), T(-v.s)));
    }
  }
  P.tw.alloc(tw.size() * sizeof(C<T>));
  VP_HIP(hipMemcpy(P.tw.p, tw.data(), P.tw.bytes, hipMemcpyHostToDevice));
  P.work.alloc(size_t(3) * g.V * sizeof(C<T>));
  if (g.level == 1 || g.qlay) P.work2.alloc(size_t(3) * g.V * sizeof(C<T>));
  P.xs.alloc(size_t(std::max<int64_t>(p->nbox, 1)) * sizeof(C<T>));
  P.ys.alloc(size_t(std::max<int64_t>(p->nbox, 1)) * sizeof(C<T>));
  const C<T>* bp = P.blocks.template as<C<T>>();
  P.blockptr.alloc(sizeof(bp));
  VP_HIP(hipMemcpy(P.blockptr.p, &bp, sizeof(bp), hipMemcpyHostToDevice));
}

// Builds the FP32 block copy from the FP64 blocks, scaled by scale32.
// Load window of the binary16 block product, timed once per (device, block
// geometry) in the process and shared by every preconditioner of that shape.
struct QuKey {
  int device, L, M, N, Hx, Hy, pct, qlay;
  bool operator<(const QuKey& o) const {
    return std::tie(device, L, M, N, Hx, Hy, pct, qlay) < std::tie(o.device, o.L, o.M, o.N, o.Hx, o.Hy, o.pct, o.qlay);
  }
};
static std::mutex g_qu_mu;
static std::map<QuKey, int> g_qu;

static int quadmul_window(vp_prec_s* p) {
  const PcGeom& g = p->g;
  const QuKey key{p->ctx->device, g.L, g.M, g.N, g.Hx, g.Hy, g.pct, g.qlay};
  {
    std::lock_guard<std::mutex> lk(g_qu_mu);
    auto it = g_qu.find(key);
    if (it != g_qu.end()) return it->second;
  }
  const PcArgs<float> a = prec_args<float>(p, nullptr, nullptr);
  const int qu = tune_quadmul_window(g, a, p->ctx->stream);
  std::lock_guard<std::mutex> lk(g_qu_mu);
  g_qu.emplace(key, qu);
  return qu;
}

static void prec_make_fp32(vp_prec_s* p, double scale32) {
  const int64_t nb = int64_t(p->p64.blocks.bytes / sizeof(double2));
  p->p32.blocks.alloc(size_t(nb) * sizeof(float2));
  convert_c128_to_c64(p->p64.blocks.as<double2>(), p->p32.blocks.as<float2>(), nb, scale32, p->ctx->stream);
  VP_HIP(hipStreamSynchronize(p->ctx->stream));
  p->scale32 = scale32;
  prec_alloc_prec<float>(p, p->p32);
  // quadrant form: the FP32 apply streams binary16 blocks (VP_PC_HALF=0 keeps FP32)
  const char* eh = std::getenv("VP_PC_HALF");
  p->p32.hblocks.release();
  if (p->g.level == 2 && p->g.pquad && !(eh && eh[0] == '0')) {
    make_half_blocks(p->p32, p->ctx->stream);
    p->g.qu = quadmul_window(p);
  }
}

// blocks_dev: FP64 blocks in (kx, ky)-major order, already inverted; scaled here by 1/(L*M)
static void prec_finish(vp_prec_s* p, int precisions) {
  const PcGeom& g = p->g;
  const int64_t nb = int64_t(p->p64.blocks.bytes / sizeof(double2));
  const double sc = 1.0 / (double(g.L) * (g.level == 2 ? g.M : 1));
  hipLaunchKernelGGL(k_scale_c128, dim3(int(std::min<int64_t>((nb + 255) / 256, 8192))), dim3(256), 0,
                     p->ctx->stream, p->p64.blocks.as<double2>(), nb, sc);
  VP_HIP(hipGetLastError());
  VP_HIP(hipStreamSynchronize(p->ctx->stream));
  prec_alloc_prec<double>(p, p->p64);
  if (precisions & VP_FP32) prec_make_fp32(p, 1.0);
}

__global__ void k_scale_c128(double2* a, int64_t n, double s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    a[i] = make_double2(a[i].x * s, a[i].y * s);
}

// An embedded apply first copies x to y (unknowns outside the box pass
// through), except for the later parts of a composite (the first part made
// the copy) and for a box holding every unknown of the outer vector: the
// embedding is then a permutation and the x inverse pass writes every entry.
static int skip_outer_copy(const vp_prec_s* q, size_t part) {
  return (part > 0 || (q->g.n_outer > 0 && q->nbox == q->g.n_outer)) ? 1 : 0;
}

template <typename T>
static void prec_apply_dev(vp_prec_s* p, const C<T>* x, C<T>* y) {
  const auto leaves = prec_leaves(p);
  for (size_t k = 0; k < leaves.size(); ++k) {
    vp_prec_s* q = leaves[k];
    PrecPrec<T>& P = pprec_of<T>(q);
    VP_REQUIRE(P.present, ERR_INVALID, "preconditioner lacks this precision");
    PcArgs<T> a = prec_args<T>(q, x, y);
    a.skip_copy = skip_outer_copy(q, k);
    launch_prec<T>(q->g, a, 1, p->ctx->stream);
  }
}

// ----------------------------------------------------------------------------
// Problem sets over operators (and preconditioners)
// ----------------------------------------------------------------------------
template <typename T>
void make_problem_set(ProblemSet<T>& ps, vp_op_s* const* ops, vp_prec_s* const* precs, int np, bool private_work) {
  vp_op_s* o0 = ops[0];
  OpPrec<T>& P0 = prec_of<T>(o0);
  VP_REQUIRE(P0.present, ERR_INVALID, "operator lacks the precision this solve needs");
  ps.nprob = np;
  ps.n = o0->n;
  ps.stream = o0->ctx->stream;
  ps.recycle_width = o0->ctx->recycle_width;
  ps.g = P0.g;
  ps.plan = P0.plan;
  ps.twx = P0.twx();
  ps.twy = P0.twy();
  ps.twz = P0.twz();
  std::vector<ProbPtrs<T>> h(np);
  const size_t wbytes = size_t(3) * P0.g.x.P * o0->Mb * o0->Nb * sizeof(C<T>);
  const size_t w2bytes = P0.plan.variant == 1 ? size_t(3) * P0.g.x.P * P0.g.y.P * o0->Nb * sizeof(C<T>) : 0;
  const bool own = np > 1 || private_work;
  if (own) {
    ps.works.alloc(wbytes * np);
    ps.works2.alloc(w2bytes * np);
  }
  for (int p = 0; p < np; ++p) {
    vp_op_s* op = ops[p];
    VP_REQUIRE(op && op->L == o0->L && op->M == o0->M && op->N == o0->N && op->n == o0->n &&
                   op->structured == o0->structured &&
                   op->l0 == o0->l0 && op->m0 == o0->m0 && op->n0 == o0->n0 && op->Lb == o0->Lb &&
                   op->Mb == o0->Mb && op->Nb == o0->Nb,
               ERR_INVALID, "batched operators must share one geometry");
    OpPrec<T>& P = prec_of<T>(op);
    VP_REQUIRE(P.present, ERR_INVALID, "operator lacks the precision this solve needs");
    // The plan is a function of (grid, FFT sizes, precision, variant, device).
    VP_REQUIRE(P.plan.variant == P0.plan.variant && P.plan.fsym == P0.plan.fsym && P.g.x.P == P0.g.x.P &&
                   P.g.y.P == P0.g.y.P && P.g.z.P == P0.g.z.P && P.g.y.R == P0.g.y.R && P.g.z.R == P0.g.z.R,
               ERR_INVALID, "batched operators differ in plan");
    ProbPtrs<T> pp{};
    pp.work = own ? reinterpret_cast<C<T>*>(static_cast<char*>(ps.works.p) + wbytes * p)
                  : P.work.template as<C<T>>();
    if (w2bytes)
      pp.work2 = own ? reinterpret_cast<C<T>*>(static_cast<char*>(ps.works2.p) + w2bytes * p)
                     : P.work2.template as<C<T>>();
    pp.fN = P.fN.template as<C<T>>();
    pp.pos = op->pos.as<int>();
    pp.mc = P.mc.template as<C<T>>();
    pp.mr = P.mr.p ? P.mr.template as<C<T>>() : nullptr;
    pp.gram = P.gram;
    h[p] = pp;
  }
  ps.probs.alloc(sizeof(ProbPtrs<T>) * np);
  VP_HIP(hipMemcpy(ps.probs.p, h.data(), ps.probs.bytes, hipMemcpyHostToDevice));
  ps.has_prec = precs && precs[0];
  if (ps.has_prec) {
    // every problem's preconditioner has the same parts (one, or the parts
    // of a composite); part k of all problems runs as one batched launch
    const auto leaves0 = prec_leaves(precs[0]);
    const size_t K = leaves0.size();
    std::vector<std::vector<vp_prec_s*>> lv(np);
    for (int p = 0; p < np; ++p) {
      lv[p] = prec_leaves(precs[p]);
      VP_REQUIRE(lv[p].size() == K, ERR_INVALID, "batched preconditioners must have the same parts");
    }
    ps.pparts.clear();
    for (size_t k = 0; k < K; ++k) {
      auto part = std::make_unique<typename ProblemSet<T>::PrecPart>();
      vp_prec_s* q0 = leaves0[k];
      std::vector<const C<T>*> bps(np);
      std::vector<const uint32_t*> hps(np, nullptr);
      std::vector<float> hsc(np, 1.f);
      bool half = true;
      for (int p = 0; p < np; ++p) {
        vp_prec_s* q = lv[p][k];
        VP_REQUIRE(q && q->g.L == q0->g.L && q->g.M == q0->g.M && q->g.N == q0->g.N && q->nbox == q0->nbox &&
                       q->g.level == q0->g.level && q->g.n_outer == q0->g.n_outer && q->g.bsym == q0->g.bsym,
                   ERR_INVALID, "batched preconditioners must share one geometry");
        // the batched launch reads problem 0's voxel map and embedding for all
        VP_REQUIRE(q->pos_hash == q0->pos_hash && q->sub_hash == q0->sub_hash && q->sub.bytes == q0->sub.bytes,
                   ERR_INVALID, "batched preconditioners must share one voxel map and embedding");
        if (std::is_same<T, float>::value && (!q->p32.present || q->scale32 != ops[p]->scale32))
          prec_make_fp32(q, ops[p]->scale32);
        PrecPrec<T>& Q = pprec_of<T>(q);
        VP_REQUIRE(Q.present, ERR_INVALID, "preconditioner lacks this precision");
        bps[p] = Q.blocks.template as<C<T>>();
        half = half && Q.hblocks.p != nullptr;
        if (Q.hblocks.p) {
          hps[p] = Q.hblocks.template as<uint32_t>();
          hsc[p] = Q.hscale;
        }
      }
      const int64_t outer = q0->g.n_outer ? q0->g.n_outer : q0->nbox;
      VP_REQUIRE(outer == ps.n, ERR_INVALID, "preconditioner size does not match the operator");
      part->pg = q0->g;
      part->pblocks.alloc(sizeof(const C<T>*) * np);
      VP_HIP(hipMemcpy(part->pblocks.p, bps.data(), part->pblocks.bytes, hipMemcpyHostToDevice));
      const int64_t wst = 3 * q0->g.V;
      part->pwork.alloc(size_t(wst) * np * sizeof(C<T>));
      if (q0->g.level == 1 || q0->g.qlay) part->pwork2.alloc(size_t(wst) * np * sizeof(C<T>));
      PcArgs<T> a = prec_args<T>(q0, nullptr, nullptr);
      a.work = part->pwork.template as<C<T>>();
      a.work2 = (q0->g.level == 1 || q0->g.qlay) ? part->pwork2.template as<C<T>>() : nullptr;
      a.wstride = wst;
      a.blocks = part->pblocks.template as<const C<T>*>();
      a.hblocks = nullptr;
      a.hscale = nullptr;
      a.skip_copy = skip_outer_copy(q0, k);
      if (half) {   // every problem's preconditioner has the binary16 copy
        part->phblocks.alloc(sizeof(const uint32_t*) * np);
        VP_HIP(hipMemcpy(part->phblocks.p, hps.data(), part->phblocks.bytes, hipMemcpyHostToDevice));
        part->phscale.alloc(sizeof(float) * np);
        VP_HIP(hipMemcpy(part->phscale.p, hsc.data(), part->phscale.bytes, hipMemcpyHostToDevice));
        a.hblocks = part->phblocks.template as<const uint32_t*>();
        a.hscale = part->phscale.template as<float>();
      }
      part->pa = a;
      ps.pparts.push_back(std::move(part));
    }
  }
}

template void make_problem_set<double>(ProblemSet<double>&, vp_op_s* const*, vp_prec_s* const*, int, bool);
template void make_problem_set<float>(ProblemSet<float>&, vp_op_s* const*, vp_prec_s* const*, int, bool);

static void fill_info(vp_gmres_info* info, const GmresResult& r, int p) {
  if (!info) return;
  info->flag = r.flag[p];
  info->iter_inner = r.iter_inner[p];
  info->iter_outer = r.iter_outer[p];
  info->total_iters = r.total[p];
  info->relres = r.relres[p];
  info->resvec_len = 0;
  if (info->resvec && info->resvec_cap > 0) {
    const int cnt = std::min<int>(info->resvec_cap, int(r.resvec[p].size()));
    for (int i = 0; i < cnt; ++i) info->resvec[i] = r.resvec[p][i];
    info->resvec_len = cnt;
  }
}

// Mixed precision: FP64 refinement around FP32 GMRES for a problem set.
struct IrResult {
  std::vector<int> flag, steps, inner;
  std::vector<double> relres;
  std::vector<std::vector<double>> hist;
};

// One refinement round over a problem set (b, x: np device vectors of n).
// For every problem not yet done: r = b - A x in FP64, relres = ||r||/||b||;
// the problem is done when relres <= tol (flag 0), when relres is not
// finite or after max_outer steps (flag 1); otherwise one inner FP32 GMRES
// solve of (A/s) d = r/||r|| and x += d*||r||/s.  The inner solve reduces
// its residual by what this step needs to reach tol (need = tol/relres/2,
// factor 2 margin for the FP32 residual gap) when that is looser than
// inner_tol or within a factor 8 below it (one slightly deeper inner solve
// instead of another refinement step), else by inner_tol.  With rec, the
// first inner solve leaves its last cycle's Krylov space as the recycled
// space of the later ones.  Problems already done do no work.  Returns true
// while a problem is still running.
bool ir_step_set(ProblemSet<double>& p64, ProblemSet<float>& p32, const double2* b, double2* x, double scale32,
                 double tol, int restart, double inner_tol, int inner_maxit, int max_outer, int side,
                 std::vector<IrProblem>& st, Recycle* rec) {
  const int np = p64.nprob;
  const int64_t n = p64.n;
  hipStream_t s = p64.stream;
  DevBuf ax(size_t(n) * np * sizeof(double2)), r(size_t(n) * np * sizeof(double2));
  DevBuf bn(np * sizeof(double)), rn(np * sizeof(double)), act(np * sizeof(int));
  std::vector<int> hact(np);
  for (int p = 0; p < np; ++p) hact[p] = st[p].done ? 0 : 1;
  VP_HIP(hipMemcpyAsync(act.p, hact.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
  // r = b - A x  (FP64); before any step x = 0 and r = b
  bool all_first = true;
  for (int p = 0; p < np; ++p) all_first &= !hact[p] || st[p].steps == 0;
  if (all_first) {
    VP_HIP(hipMemcpyAsync(r.p, b, size_t(n) * np * sizeof(double2), hipMemcpyDeviceToDevice, s));
  } else {
    ps_apply_A<double>(p64, x, n, ax.as<double2>(), n, act.as<int>());
    dev_residual_d(b, ax.as<double2>(), r.as<double2>(), n, np, s);
  }
  dev_norms<double>(r.as<double2>(), n, np, n, rn.as<double>(), s);
  dev_norms<double>(b, n, np, n, bn.as<double>(), s);
  std::vector<double> hbn(np), hrn(np);
  VP_HIP(hipMemcpyAsync(hrn.data(), rn.p, np * sizeof(double), hipMemcpyDeviceToHost, s));
  VP_HIP(hipMemcpyAsync(hbn.data(), bn.p, np * sizeof(double), hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  bool any = false;
  for (int p = 0; p < np; ++p) {
    if (!hact[p]) continue;
    IrProblem& q = st[p];
    q.relres = relres_of(hrn[p], hbn[p]);
    q.hist.push_back(q.relres);
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
  if (progress_enabled()) {
    for (int p = 0; p < np; ++p)
      std::fprintf(stderr, "[vp] refinement: problem %d step %d, FP64 relres %.3e, inner iterations %d%s\n", p,
                   st[p].steps, st[p].relres, st[p].inner, st[p].done ? " (done)" : "");
    std::fflush(stderr);
  }
  if (!any) return false;
  VP_HIP(hipMemcpyAsync(act.p, hact.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
  DevBuf r32(size_t(n) * np * sizeof(float2)), d32(size_t(n) * np * sizeof(float2));
  dev_scale_to_f(r.as<double2>(), rn.as<double>(), r32.as<float2>(), n, np, 1.0, s);
  GmresParams prm;
  prm.restart = restart;
  prm.tol = inner_tol;
  prm.maxit = inner_maxit;
  prm.side = side;
  prm.final_residual = false;
  prm.active_in = hact;
  prm.tol_per.assign(np, inner_tol);
  for (int p = 0; p < np; ++p)
    if (hact[p] && st[p].relres > 0) {
      const double need = std::min(0.5, 0.5 * tol / st[p].relres);
      prm.tol_per[p] = need >= inner_tol / 8 ? need : inner_tol;
    }
  if (rec) {
    // the first inner solve leaves its Krylov space as the recycled space of
    // the later ones (per problem where the members bring their own spaces)
    prm.rec = rec;
    if (rec->want.empty() ? rec->kc == 0
                          : std::any_of(rec->want.begin(), rec->want.end(), [](int v) { return v != 0; }))
      rec->build = true;
  }
  GmresResult gr;
  gmres_device<float>(p32, r32.as<float2>(), d32.as<float2>(), prm, gr);
  for (int p = 0; p < np; ++p)
    if (hact[p]) {
      st[p].inner += gr.total[p];
      st[p].steps += 1;
    }
  // x += d * ||r|| / s
  dev_update_from_f(x, d32.as<float2>(), rn.as<double>(), n, np, 1.0 / scale32, act.as<int>(), s);
  VP_HIP(hipStreamSynchronize(s));
  return true;
}

// outer_m = 0: one correction per refinement step (ir_step_set); > 0:
// flexible refinement in cycles of outer_m steps (fgmres_ir_set)
static void solve_ir_sets(ProblemSet<double>& p64, ProblemSet<float>& p32, const double2* b, double2* x,
                          double scale32, double tol, int restart, double inner_tol, int inner_maxit,
                          int max_outer, int side, int outer_m, IrResult& out) {
  const int np = p64.nprob;
  VP_HIP(hipMemsetAsync(x, 0, size_t(p64.n) * np * sizeof(double2), p64.stream));
  std::vector<IrProblem> st(np);
  // inner solves after the first run augmented with the first one's Krylov
  // space (VP_IR_RECYCLE=0: every inner solve starts afresh)
  const char* er = std::getenv("VP_IR_RECYCLE");
  Recycle rec;
  Recycle* prec = nullptr;
  if (!(er && er[0] == '0')) {
    rec.alloc(p64.n, np, restart, sizeof(float2));
    prec = &rec;
  }
  if (outer_m > 0) {
    fgmres_ir_set(p64, p32, b, x, scale32, tol, restart, inner_tol, inner_maxit, max_outer, outer_m, side, st,
                  prec);
  } else {
    while (ir_step_set(p64, p32, b, x, scale32, tol, restart, inner_tol, inner_maxit, max_outer, side, st, prec)) {
    }
  }
  out.flag.resize(np);
  out.steps.resize(np);
  out.inner.resize(np);
  out.relres.resize(np);
  out.hist.resize(np);
  for (int p = 0; p < np; ++p) {
    out.flag[p] = st[p].flag;
    out.steps[p] = st[p].steps;
    out.inner[p] = st[p].inner;
    out.relres[p] = st[p].relres;
    out.hist[p] = st[p].hist;
  }
}

}  // namespace vp

extern "C" {

int vp_prec_create_chan2(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* blocks, const int64_t* idx,
                         int64_t n, int precisions, vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(out && blocks, ERR_INVALID, "out/blocks is NULL");
  vp_prec_s* p = prec_begin(ctx, L, M, N, idx, n, 2);
  try {
    const int64_t n3 = 3 * N;
    const int64_t bs = n3 * n3;
    std::vector<double2> h(size_t(L * M * bs));
    for (int64_t j = 0; j < M; ++j)        // cell {i,j} at (i + L*j) -> (i*M + j)
      for (int64_t i = 0; i < L; ++i) {
        const double* src = blocks + 2 * (i + L * j) * bs;
        double2* dst = h.data() + (i * M + j) * bs;
        for (int64_t e = 0; e < bs; ++e) dst[e] = make_double2(src[2 * e], src[2 * e + 1]);
      }
    p->p64.blocks.alloc(h.size() * sizeof(double2));
    VP_HIP(hipMemcpy(p->p64.blocks.p, h.data(), p->p64.blocks.bytes, hipMemcpyHostToDevice));
    prec_finish(p, precisions);
  } catch (...) {
    delete p;
    throw;
  }
  *out = p;
  VP_API_END
}

int vp_prec_create_chan1(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* blocks, const int64_t* idx,
                         int64_t n, int precisions, vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(out && blocks, ERR_INVALID, "out/blocks is NULL");
  vp_prec_s* p = prec_begin(ctx, L, M, N, idx, n, 1);
  try {
    const int64_t n3 = 3 * M * N;
    const size_t cnt = size_t(L * n3 * n3);
    p->p64.blocks.alloc(cnt * sizeof(double2));
    VP_HIP(hipMemcpy(p->p64.blocks.p, blocks, cnt * sizeof(double2), hipMemcpyHostToDevice));
    prec_finish(p, precisions);
  } catch (...) {
    delete p;
    throw;
  }
  *out = p;
  VP_API_END
}

// Device-built 2-level blocks are kept as the (kx <= L/2, ky <= M/2)
// quadrant when the tensor has the parity zeros that make the Chan
// circulants symmetric (VP_PC_SYMMETRY=0 keeps all L*M blocks).
static void level2_block_range(vp_prec_s* p, const double2* dG) {
  const char* e = std::getenv("VP_PC_SYMMETRY");
  PcGeom& g = p->g;
  g.bsym = (!(e && std::strcmp(e, "0") == 0) && toeplitz_parity_zero(dG, g.L, g.M, g.N, p->ctx->stream)) ? 1 : 0;
  g.Hx = g.bsym ? g.L / 2 + 1 : g.L;
  g.Hy = g.bsym ? g.M / 2 + 1 : g.M;
  // quadrant blocks: one read per stored block for its four mirrored
  // columns (VP_PC_QUAD=0 keeps the per-plane kernel)
  const char* eq = std::getenv("VP_PC_QUAD");
  g.pquad = (g.bsym && 3 * g.N <= 255 && g.Brow > 0 && !(eq && eq[0] == '0')) ? 1 : 0;
  // pencil-major spectra with compile-time x and row passes where both
  // lengths have them in both precisions (VP_PC_CT=0 keeps plane-major)
  const char* ec = std::getenv("VP_PC_CT");
  const int lds_max = max_lds_bytes(p->ctx->device);
  g.pct = (g.pquad && !(ec && ec[0] == '0') && xct_threads(g.L, false) > 0 && xct_threads(g.L, true) > 0 &&
           pcrow_supported(g.M) && (size_t(g.L) + 64 + (size_t(g.L) + 63) / 64) * 16 <= size_t(lds_max))
              ? 1
              : 0;
  // q-contiguous spectra for the block product (VP_PC_QLAY=0: gather the
  // columns from the pencil-major spectra)
  const char* eqc = std::getenv("VP_PC_QLAY");
  g.qlay = (g.pct && !(eqc && eqc[0] == '0')) ? 1 : 0;
}

static int prec_level2_common(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double2* dG, double mc_re,
                              double mc_im, double mr0, double gram, const int64_t* idx, int64_t n,
                              int precisions, vp_prec* out) {
  vp_prec_s* p = prec_begin(ctx, L, M, N, idx, n, 2);
  try {
    const int64_t n3 = 3 * N;
    level2_block_range(p, dG);
    const int64_t nb = int64_t(p->g.Hx) * p->g.Hy;
    p->p64.blocks.alloc(size_t(nb * n3 * n3) * sizeof(double2));
    build_level2_blocks(dG, int(L), int(M), int(N), make_double2(mc_re, mc_im), gram * mr0,
                        p->p64.blocks.as<double2>(), p->g.Hx, p->g.Hy, ctx->stream);
    invert_blocks(p->p64.blocks.as<double2>(), nb, int(n3), ctx->stream);
    prec_finish(p, precisions);
  } catch (...) {
    delete p;
    throw;
  }
  *out = p;
  return VP_OK;
}

// G_k(d) *= Mc(d) over the L x M x N offsets of each of the 6 components: the
// contrast field folded into the Toeplitz tensor, level_2_parallel_func.m:5,
// 24, 36, 47, 61, 74 (-Mc.*opToeplitz(:,:,:,k) with an array Mc).
__global__ void k_fold_contrast(double2* __restrict__ G, const double2* __restrict__ mc, int64_t V) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < 6 * V; e += int64_t(gridDim.x) * blockDim.x) {
    const double2 a = G[e], b = mc[e % V];
    G[e] = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
}

static void fold_contrast(double2* dG, const double* Mc, int64_t V, hipStream_t s) {
  DevBuf dm(size_t(V) * sizeof(double2));
  VP_HIP(hipMemcpyAsync(dm.p, Mc, dm.bytes, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fold_contrast, dim3(unsigned(std::min<int64_t>((6 * V + 255) / 256, 8192))), dim3(256), 0, s,
                     dG, dm.as<double2>(), V);
  VP_HIP(hipGetLastError());
  VP_HIP(hipStreamSynchronize(s));
}

int vp_prec_create_level2(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* G, double mc_re, double mc_im,
                          double mr0, double gram, const int64_t* idx, int64_t n, int precisions, vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && out && G, ERR_INVALID, "ctx/out/G is NULL");
  VP_HIP(hipSetDevice(ctx->device));
  DevBuf dG(size_t(L * M * N) * 6 * sizeof(double2));
  VP_HIP(hipMemcpy(dG.p, G, dG.bytes, hipMemcpyHostToDevice));
  prec_level2_common(ctx, L, M, N, dG.as<double2>(), mc_re, mc_im, mr0, gram, idx, n, precisions, out);
  VP_API_END
}

int vp_prec_create_level2_assembled(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko, double damping,
                                    double mc_re,
                                    double mc_im, double mr0, const int64_t* idx, int64_t n, int precisions,
                                    vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && out && dx > 0 && damping >= 0, ERR_INVALID, "bad arguments");
  VP_HIP(hipSetDevice(ctx->device));
  DevBuf dG(size_t(L * M * N) * 6 * sizeof(double2));
  assemble_toeplitz_dev(int(L), int(M), int(N), dx, ko, dG.as<double2>(), ctx->stream, -damping * ko);
  prec_level2_common(ctx, L, M, N, dG.as<double2>(), mc_re, mc_im, mr0, dx * dx * dx, idx, n, precisions, out);
  VP_API_END
}

int vp_prec_create_level2_field(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* G, const double* Mc,
                                double mr0, double gram, const int64_t* idx, int64_t n, int precisions, vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && out && G && Mc, ERR_INVALID, "ctx/out/G/Mc is NULL");
  VP_REQUIRE(L > 0 && M > 0 && N > 0, ERR_INVALID, "bad box");
  VP_HIP(hipSetDevice(ctx->device));
  const int64_t V = L * M * N;
  DevBuf dG(size_t(V) * 6 * sizeof(double2));
  VP_HIP(hipMemcpy(dG.p, G, dG.bytes, hipMemcpyHostToDevice));
  fold_contrast(dG.as<double2>(), Mc, V, ctx->stream);
  prec_level2_common(ctx, L, M, N, dG.as<double2>(), 1.0, 0.0, mr0, gram, idx, n, precisions, out);
  VP_API_END
}

int vp_prec_create_level2_assembled_field(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko,
                                          double damping, const double* Mc, double mr0, const int64_t* idx, int64_t n,
                                          int precisions, vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && out && Mc && dx > 0 && damping >= 0, ERR_INVALID, "bad arguments");
  VP_REQUIRE(L > 0 && M > 0 && N > 0, ERR_INVALID, "bad box");
  VP_HIP(hipSetDevice(ctx->device));
  const int64_t V = L * M * N;
  DevBuf dG(size_t(V) * 6 * sizeof(double2));
  assemble_toeplitz_dev(int(L), int(M), int(N), dx, ko, dG.as<double2>(), ctx->stream, -damping * ko);
  fold_contrast(dG.as<double2>(), Mc, V, ctx->stream);
  prec_level2_common(ctx, L, M, N, dG.as<double2>(), 1.0, 0.0, mr0, dx * dx * dx, idx, n, precisions, out);
  VP_API_END
}

// 1-level setup (level_1_parallel_func.m): blocks of the folded tensor
// (-mc G, Gram*Mr(1) at the origin) or, face != nullptr, Gram*I -
// diag(face3) circ_N of the bare tensor (level_1_parallel_func_N.m with
// Bragg_MOR_script.m:61-72); all L blocks, inverted on the device.
static int prec_level1_common(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double2* dG, double2 mc,
                              double diag, const double* face, double gram, const int64_t* idx, int64_t n,
                              int precisions, vp_prec* out) {
  VP_REQUIRE(L > 0 && M > 0 && N > 0, ERR_INVALID, "bad box");
  const int64_t n3 = 3 * M * N;
  VP_REQUIRE(n3 <= 16384, ERR_UNSUPPORTED, "1-level blocks wider than 16384 rows");
  vp_prec_s* p = prec_begin(ctx, L, M, N, idx, n, 1);
  try {
    p->p64.blocks.alloc(size_t(L) * n3 * n3 * sizeof(double2));
    DevBuf dface;
    if (face) {
      dface.alloc(size_t(M * N) * sizeof(double2));
      VP_HIP(hipMemcpy(dface.p, face, dface.bytes, hipMemcpyHostToDevice));
    }
    build_level1_blocks(dG, int(L), int(M), int(N), mc, diag, p->p64.blocks.as<double2>(), ctx->stream,
                        face ? dface.as<double2>() : nullptr, gram);
    invert_blocks_any(p->p64.blocks.as<double2>(), L, int(n3), ctx->stream);
    prec_finish(p, precisions);
  } catch (...) {
    delete p;
    throw;
  }
  *out = p;
  return VP_OK;
}

int vp_prec_create_level1(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* G, double mc_re, double mc_im,
                          const double* Mc, double mr0, double gram, const int64_t* idx, int64_t n, int precisions,
                          vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && out && G, ERR_INVALID, "ctx/out/G is NULL");
  VP_REQUIRE(L > 0 && M > 0 && N > 0, ERR_INVALID, "bad box");
  VP_HIP(hipSetDevice(ctx->device));
  const int64_t V = L * M * N;
  DevBuf dG(size_t(V) * 6 * sizeof(double2));
  VP_HIP(hipMemcpy(dG.p, G, dG.bytes, hipMemcpyHostToDevice));
  double2 mc = make_double2(mc_re, mc_im);
  if (Mc) {
    fold_contrast(dG.as<double2>(), Mc, V, ctx->stream);
    mc = make_double2(1.0, 0.0);
  }
  prec_level1_common(ctx, L, M, N, dG.as<double2>(), mc, gram * mr0, nullptr, 0.0, idx, n, precisions, out);
  VP_API_END
}

int vp_prec_create_level1_assembled(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko,
                                    double damping, double mc_re, double mc_im, const double* Mc, double mr0,
                                    const int64_t* idx, int64_t n, int precisions, vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && out && dx > 0 && damping >= 0, ERR_INVALID, "bad arguments");
  VP_REQUIRE(L > 0 && M > 0 && N > 0, ERR_INVALID, "bad box");
  VP_HIP(hipSetDevice(ctx->device));
  const int64_t V = L * M * N;
  DevBuf dG(size_t(V) * 6 * sizeof(double2));
  assemble_toeplitz_dev(int(L), int(M), int(N), dx, ko, dG.as<double2>(), ctx->stream, -damping * ko);
  double2 mc = make_double2(mc_re, mc_im);
  if (Mc) {
    fold_contrast(dG.as<double2>(), Mc, V, ctx->stream);
    mc = make_double2(1.0, 0.0);
  }
  prec_level1_common(ctx, L, M, N, dG.as<double2>(), mc, dx * dx * dx * mr0, nullptr, 0.0, idx, n, precisions,
                     out);
  VP_API_END
}

int vp_prec_create_level1_face(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko, double damping,
                               const double* face, const int64_t* idx, int64_t n, int precisions, vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && out && face && dx > 0 && damping >= 0, ERR_INVALID, "bad arguments");
  VP_REQUIRE(L > 0 && M > 0 && N > 0, ERR_INVALID, "bad box");
  VP_HIP(hipSetDevice(ctx->device));
  DevBuf dG(size_t(L * M * N) * 6 * sizeof(double2));
  assemble_toeplitz_dev(int(L), int(M), int(N), dx, ko, dG.as<double2>(), ctx->stream, -damping * ko);
  prec_level1_common(ctx, L, M, N, dG.as<double2>(), make_double2(-1.0, 0.0), 0.0, face, dx * dx * dx, idx, n,
                     precisions, out);
  VP_API_END
}

int vp_prec_create_level2_face(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko, double damping,
                               const double* face, const int64_t* idx, int64_t n, int precisions, vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && out && face && dx > 0 && damping >= 0, ERR_INVALID, "bad arguments");
  VP_HIP(hipSetDevice(ctx->device));
  DevBuf dG(size_t(L * M * N) * 6 * sizeof(double2));
  assemble_toeplitz_dev(int(L), int(M), int(N), dx, ko, dG.as<double2>(), ctx->stream, -damping * ko);
  DevBuf dface(size_t(N) * sizeof(double2));
  VP_HIP(hipMemcpy(dface.p, face, dface.bytes, hipMemcpyHostToDevice));
  vp_prec_s* p = prec_begin(ctx, L, M, N, idx, n, 2);
  try {
    const int64_t n3 = 3 * N;
    level2_block_range(p, dG.as<double2>());
    const int64_t nb = int64_t(p->g.Hx) * p->g.Hy;
    p->p64.blocks.alloc(size_t(nb * n3 * n3) * sizeof(double2));
    build_level2_blocks(dG.as<double2>(), int(L), int(M), int(N), make_double2(-1.0, 0.0), 0.0,
                        p->p64.blocks.as<double2>(), p->g.Hx, p->g.Hy, ctx->stream, dface.as<double2>(),
                        dx * dx * dx);
    invert_blocks(p->p64.blocks.as<double2>(), nb, int(n3), ctx->stream);
    prec_finish(p, precisions);
  } catch (...) {
    delete p;
    throw;
  }
  *out = p;
  VP_API_END
}

int vp_prec_set_embedding(vp_prec p, int64_t n_outer, const int64_t* sub_of_op, int64_t nsub) {
  VP_API_BEGIN
  VP_REQUIRE(p && sub_of_op, ERR_INVALID, "p/sub is NULL");
  VP_REQUIRE(p->parts.empty(), ERR_INVALID, "a composite takes its parts' embeddings");
  VP_REQUIRE(nsub == p->nbox, ERR_INVALID, "embedding length must equal the number of box unknowns");
  VP_REQUIRE(n_outer >= nsub, ERR_INVALID, "outer vector shorter than the box");
  std::vector<int> h(static_cast<size_t>(nsub));
  std::vector<char> seen(static_cast<size_t>(n_outer), 0);
  for (int64_t i = 0; i < nsub; ++i) {
    VP_REQUIRE(sub_of_op[i] >= 1 && sub_of_op[i] <= n_outer, ERR_INVALID, "embedding index out of range");
    VP_REQUIRE(!seen[size_t(sub_of_op[i] - 1)], ERR_INVALID, "embedding index repeated");
    seen[size_t(sub_of_op[i] - 1)] = 1;
    h[size_t(i)] = int(sub_of_op[i] - 1);
  }
  VP_HIP(hipSetDevice(p->ctx->device));
  p->sub.alloc(h.size() * sizeof(int));
  VP_HIP(hipMemcpy(p->sub.p, h.data(), p->sub.bytes, hipMemcpyHostToDevice));
  p->sub_hash = map_hash(h);
  p->sub_host = h;
  p->g.n_outer = n_outer;
  p->p64.xs.alloc(size_t(n_outer) * sizeof(double2));
  p->p64.ys.alloc(size_t(n_outer) * sizeof(double2));
  if (p->p32.present) {
    p->p32.xs.alloc(size_t(n_outer) * sizeof(float2));
    p->p32.ys.alloc(size_t(n_outer) * sizeof(float2));
  }
  VP_API_END
}

int vp_prec_set_fp32_scale(vp_prec p, double scale) {
  VP_API_BEGIN
  VP_REQUIRE(p && std::isfinite(scale) && scale > 0, ERR_INVALID, "bad scale");
  VP_HIP(hipSetDevice(p->ctx->device));
  for (vp_prec_s* q : prec_leaves(p)) {
    prec_make_fp32(q, scale);
    if (q->g.n_outer) {
      q->p32.xs.alloc(size_t(q->g.n_outer) * sizeof(float2));
      q->p32.ys.alloc(size_t(q->g.n_outer) * sizeof(float2));
    }
  }
  if (!p->parts.empty()) {
    p->scale32 = scale;
    p->p32.present = true;
  }
  if (p->g.n_outer) {
    p->p32.xs.alloc(size_t(p->g.n_outer) * sizeof(float2));
    p->p32.ys.alloc(size_t(p->g.n_outer) * sizeof(float2));
  }
  VP_API_END
}

int vp_prec_apply(vp_prec p, const void* x, void* y, int mem, int prec) {
  VP_API_BEGIN
  VP_REQUIRE(p && x && y && x != y, ERR_INVALID, "p/x/y invalid (x and y must not alias)");
  VP_REQUIRE(prec == VP_FP64 || prec == VP_FP32, ERR_INVALID, "prec must be VP_FP64 or VP_FP32");
  VP_HIP(hipSetDevice(p->ctx->device));
  hipStream_t s = p->ctx->stream;
  const int64_t n = p->g.n_outer ? p->g.n_outer : p->nbox;
  if (mem == VP_MEM_DEVICE) {
    if (prec == VP_FP64) prec_apply_dev<double>(p, static_cast<const double2*>(x), static_cast<double2*>(y));
    else prec_apply_dev<float>(p, static_cast<const float2*>(x), static_cast<float2*>(y));
  } else if (mem == VP_MEM_HOST) {
    if (prec == VP_FP64) {
      VP_HIP(hipMemcpyAsync(p->p64.xs.p, x, n * sizeof(double2), hipMemcpyHostToDevice, s));
      prec_apply_dev<double>(p, p->p64.xs.as<double2>(), p->p64.ys.as<double2>());
      VP_HIP(hipMemcpyAsync(y, p->p64.ys.p, n * sizeof(double2), hipMemcpyDeviceToHost, s));
    } else {
      VP_REQUIRE(p->p32.present, ERR_INVALID, "preconditioner has no FP32 copy");
      DevBuf t(n * sizeof(double2));
      VP_HIP(hipMemcpyAsync(t.p, x, n * sizeof(double2), hipMemcpyHostToDevice, s));
      convert_c128_to_c64(t.as<double2>(), p->p32.xs.as<float2>(), n, 1.0, s);
      prec_apply_dev<float>(p, p->p32.xs.as<float2>(), p->p32.ys.as<float2>());
      convert_c64_to_c128(p->p32.ys.as<float2>(), t.as<double2>(), n, 1.0 / p->scale32, s);
      VP_HIP(hipMemcpyAsync(y, t.p, n * sizeof(double2), hipMemcpyDeviceToHost, s));
      VP_HIP(hipStreamSynchronize(s));
    }
    VP_HIP(hipStreamSynchronize(s));
  } else {
    throw VpError{ERR_INVALID, "mem must be VP_MEM_HOST or VP_MEM_DEVICE"};
  }
  VP_API_END
}

int vp_prec_create_composite(vp_ctx ctx, const vp_prec* parts, int nparts, vp_prec* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && parts && out && nparts > 0, ERR_INVALID, "bad composite arguments");
  int64_t n_outer = -1, nbox = 0;
  bool has64 = true, has32 = true;
  for (int k = 0; k < nparts; ++k) {
    vp_prec_s* q = parts[k];
    VP_REQUIRE(q && q->ctx == ctx, ERR_INVALID, "composite parts must be preconditioners of this context");
    VP_REQUIRE(q->parts.empty(), ERR_INVALID, "a composite part must not itself be a composite");
    VP_REQUIRE(q->g.n_outer > 0 && !q->sub_host.empty(), ERR_INVALID,
               "every composite part needs an embedding (vp_prec_set_embedding)");
    VP_REQUIRE(n_outer < 0 || q->g.n_outer == n_outer, ERR_INVALID,
               "composite parts must embed into one outer vector");
    n_outer = q->g.n_outer;
    nbox += q->nbox;
    has64 = has64 && q->p64.present;
    has32 = has32 && q->p32.present;
    for (int j = 0; j < k; ++j) VP_REQUIRE(parts[j] != q, ERR_INVALID, "a part appears twice");
  }
  VP_REQUIRE(has64 || has32, ERR_INVALID, "the parts share no precision");
  // disjoint index sets: block Jacobi
  std::vector<char> seen(size_t(n_outer), 0);
  for (int k = 0; k < nparts; ++k)
    for (int i : parts[k]->sub_host) {
      VP_REQUIRE(!seen[size_t(i)], ERR_INVALID, "composite parts must cover disjoint unknowns");
      seen[size_t(i)] = 1;
    }
  VP_HIP(hipSetDevice(ctx->device));
  auto* p = new vp_prec_s();
  try {
    p->ctx = ctx;
    p->parts.assign(parts, parts + nparts);
    p->g = parts[0]->g;          // outer length and a representative geometry
    p->g.n_outer = n_outer;
    p->nbox = nbox;
    p->scale32 = parts[0]->scale32;
    p->p64.present = has64;
    p->p32.present = has32;
    uint64_t h = 1469598103934665603ull;
    for (int k = 0; k < nparts; ++k) h = (h ^ parts[k]->pos_hash ^ (parts[k]->sub_hash << 1)) * 1099511628211ull;
    p->pos_hash = h;
    p->sub_hash = h;
    if (has64) {
      p->p64.xs.alloc(size_t(n_outer) * sizeof(double2));
      p->p64.ys.alloc(size_t(n_outer) * sizeof(double2));
    }
    if (has32) {
      p->p32.xs.alloc(size_t(n_outer) * sizeof(float2));
      p->p32.ys.alloc(size_t(n_outer) * sizeof(float2));
    }
  } catch (...) {
    delete p;
    throw;
  }
  *out = p;
  VP_API_END
}

int vp_prec_destroy(vp_prec p) {
  VP_API_BEGIN
  if (!p) return VP_OK;
  VP_HIP(hipSetDevice(p->ctx->device));
  (void)hipStreamSynchronize(p->ctx->stream);
  delete p;
  VP_API_END
}

int vp_gmres(vp_op op, vp_prec M, int side, const void* b, const void* x0, void* x, int mem, int prec, int restart,
             double tol, int maxit, vp_gmres_info* info) {
  VP_API_BEGIN
  VP_REQUIRE(op && b && x, ERR_INVALID, "op/b/x is NULL");
  VP_REQUIRE(restart > 0 && restart <= 1000 && maxit > 0 && tol > 0, ERR_INVALID, "bad GMRES parameters");
  VP_REQUIRE(side == VP_PREC_LEFT || side == VP_PREC_RIGHT, ERR_INVALID, "side must be LEFT or RIGHT");
  VP_REQUIRE(prec == VP_FP64 || prec == VP_FP32, ERR_INVALID, "prec must be VP_FP64 or VP_FP32");
  VP_HIP(hipSetDevice(op->ctx->device));
  hipStream_t s = op->ctx->stream;
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = op->n;
  GmresParams prm;
  prm.restart = restart;
  prm.tol = tol;
  prm.maxit = maxit;
  prm.side = side;
  prm.x0_nonzero = x0 != nullptr;
  GmresResult res;
  vp_op_s* ops[1] = {op};
  vp_prec_s* precs[1] = {M};
  auto run = [&](auto tag) {
    using T = decltype(tag);
    ProblemSet<T> ps;
    make_problem_set<T>(ps, ops, M ? precs : nullptr, 1);
    const double sc = std::is_same<T, float>::value ? 1.0 / op->scale32 : 1.0;   // b for A/scale
    if (mem == VP_MEM_DEVICE) {
      if (x0) VP_HIP(hipMemcpyAsync(x, x0, n * sizeof(C<T>), hipMemcpyDeviceToDevice, s));
      gmres_device<T>(ps, static_cast<const C<T>*>(b), static_cast<C<T>*>(x), prm, res);
    } else {
      DevBuf db(n * sizeof(C<T>)), dx(n * sizeof(C<T>)), t(n * sizeof(double2));
      VP_HIP(hipMemcpyAsync(t.p, b, n * sizeof(double2), hipMemcpyHostToDevice, s));
      if (std::is_same<T, float>::value) convert_c128_to_c64(t.as<double2>(), db.as<float2>(), n, sc, s);
      else VP_HIP(hipMemcpyAsync(db.p, t.p, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
      if (x0) {
        VP_HIP(hipMemcpyAsync(t.p, x0, n * sizeof(double2), hipMemcpyHostToDevice, s));
        if (std::is_same<T, float>::value) convert_c128_to_c64(t.as<double2>(), dx.as<float2>(), n, 1.0, s);
        else VP_HIP(hipMemcpyAsync(dx.p, t.p, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
      }
      gmres_device<T>(ps, db.as<C<T>>(), dx.as<C<T>>(), prm, res);
      if (std::is_same<T, float>::value) {
        convert_c64_to_c128(dx.as<float2>(), t.as<double2>(), n, 1.0, s);
        VP_HIP(hipMemcpyAsync(x, t.p, n * sizeof(double2), hipMemcpyDeviceToHost, s));
      } else {
        VP_HIP(hipMemcpyAsync(x, dx.p, n * sizeof(double2), hipMemcpyDeviceToHost, s));
      }
      VP_HIP(hipStreamSynchronize(s));
    }
  };
  if (prec == VP_FP64) run(double{});
  else run(float{});
  fill_info(info, res, 0);
  if (info) {
    info->ir_steps = 0;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  VP_API_END
}

static void ir_common(vp_op_s* const* ops, vp_prec_s* const* precs, int np, int side, const double2* db,
                      double2* dx, double tol, int restart, double inner_tol, int inner_maxit, int max_outer,
                      int outer_m, vp_gmres_info* infos) {
  VP_REQUIRE(restart > 0 && restart <= 1000 && inner_maxit > 0 && max_outer >= 0 && tol > 0 && inner_tol > 0,
             ERR_INVALID, "bad solve parameters");
  VP_REQUIRE(outer_m >= 0 && outer_m <= 1000, ERR_INVALID, "outer restart must be 0..1000");
  for (int p = 0; p < np; ++p)
    VP_REQUIRE(ops[p]->p64.present && ops[p]->p32.present, ERR_INVALID,
               "mixed precision needs operators created with VP_FP64|VP_FP32");
  const auto t0 = std::chrono::steady_clock::now();
  ProblemSet<double> p64;
  make_problem_set<double>(p64, ops, nullptr, np);
  ProblemSet<float> p32;
  make_problem_set<float>(p32, ops, precs, np);
  IrResult r;
  solve_ir_sets(p64, p32, db, dx, ops[0]->scale32, tol, restart, inner_tol, inner_maxit, max_outer, side, outer_m,
                r);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (infos)
    for (int p = 0; p < np; ++p) {
      vp_gmres_info* info = infos + p;
      info->flag = r.flag[p];
      info->relres = r.relres[p];
      info->ir_steps = r.steps[p];
      info->total_iters = r.inner[p];
      info->iter_inner = 0;
      info->iter_outer = r.steps[p];
      info->seconds = secs;
      info->resvec_len = 0;
      if (info->resvec && info->resvec_cap > 0) {
        const int cnt = std::min<int>(info->resvec_cap, int(r.hist[p].size()));
        for (int i = 0; i < cnt; ++i) info->resvec[i] = r.hist[p][i];
        info->resvec_len = cnt;
      }
    }
}

static void solve_one(vp_op op, vp_prec M, int side, const void* b, void* x, int mem, double tol, int restart,
                      double inner_tol, int inner_maxit, int max_outer, int outer_m, vp_gmres_info* info) {
  VP_REQUIRE(op && b && x, ERR_INVALID, "op/b/x is NULL");
  VP_REQUIRE(side == VP_PREC_LEFT || side == VP_PREC_RIGHT, ERR_INVALID, "side must be LEFT or RIGHT");
  VP_HIP(hipSetDevice(op->ctx->device));
  hipStream_t s = op->ctx->stream;
  const int64_t n = op->n;
  vp_op_s* ops[1] = {op};
  vp_prec_s* precs[1] = {M};
  if (mem == VP_MEM_DEVICE) {
    ir_common(ops, M ? precs : nullptr, 1, side, static_cast<const double2*>(b), static_cast<double2*>(x), tol,
              restart, inner_tol, inner_maxit, max_outer, outer_m, info);
  } else {
    DevBuf db(n * sizeof(double2)), dx(n * sizeof(double2));
    VP_HIP(hipMemcpyAsync(db.p, b, n * sizeof(double2), hipMemcpyHostToDevice, s));
    ir_common(ops, M ? precs : nullptr, 1, side, db.as<double2>(), dx.as<double2>(), tol, restart, inner_tol,
              inner_maxit, max_outer, outer_m, info);
    VP_HIP(hipMemcpyAsync(x, dx.p, n * sizeof(double2), hipMemcpyDeviceToHost, s));
    VP_HIP(hipStreamSynchronize(s));
  }
}

static void solve_batched(const vp_op* ops, const vp_prec* precs, int nsolve, int side, const void* b, void* x,
                          double tol, int restart, double inner_tol, int inner_maxit, int max_outer, int outer_m,
                          vp_gmres_info* infos) {
  VP_REQUIRE(ops && nsolve > 0 && b && x, ERR_INVALID, "bad batched arguments");
  VP_REQUIRE(side == VP_PREC_LEFT || side == VP_PREC_RIGHT, ERR_INVALID, "side must be LEFT or RIGHT");
  VP_REQUIRE(ops[0], ERR_INVALID, "ops[0] is NULL");
  bool anyp = false, allp = true;
  if (precs)
    for (int p = 0; p < nsolve; ++p) {
      anyp |= precs[p] != nullptr;
      allp &= precs[p] != nullptr;
    }
  VP_REQUIRE(!anyp || allp, ERR_INVALID, "either every solve has a preconditioner or none does");
  VP_HIP(hipSetDevice(ops[0]->ctx->device));
  std::vector<vp_op_s*> o(ops, ops + nsolve);
  std::vector<vp_prec_s*> q;
  if (anyp) q.assign(precs, precs + nsolve);
  ir_common(o.data(), anyp ? q.data() : nullptr, nsolve, side, static_cast<const double2*>(b),
            static_cast<double2*>(x), tol, restart, inner_tol, inner_maxit, max_outer, outer_m, infos);
}

int vp_solve_ir(vp_op op, vp_prec M, int side, const void* b, void* x, int mem, double tol, int restart,
                double inner_tol, int inner_maxit, int max_outer, vp_gmres_info* info) {
  VP_API_BEGIN
  solve_one(op, M, side, b, x, mem, tol, restart, inner_tol, inner_maxit, max_outer, 0, info);
  VP_API_END
}

int vp_solve_ir_batched(const vp_op* ops, const vp_prec* precs, int nsolve, int side, const void* b, void* x,
                        double tol, int restart, double inner_tol, int inner_maxit, int max_outer,
                        vp_gmres_info* infos) {
  VP_API_BEGIN
  solve_batched(ops, precs, nsolve, side, b, x, tol, restart, inner_tol, inner_maxit, max_outer, 0, infos);
  VP_API_END
}

int vp_solve_fgmres_ir(vp_op op, vp_prec M, int side, const void* b, void* x, int mem, double tol, int restart,
                       double inner_tol, int inner_maxit, int max_outer, int outer_restart, vp_gmres_info* info) {
  VP_API_BEGIN
  VP_REQUIRE(outer_restart > 0, ERR_INVALID, "outer_restart must be positive");
  solve_one(op, M, side, b, x, mem, tol, restart, inner_tol, inner_maxit, max_outer, outer_restart, info);
  VP_API_END
}

int vp_solve_fgmres_ir_batched(const vp_op* ops, const vp_prec* precs, int nsolve, int side, const void* b,
                               void* x, double tol, int restart, double inner_tol, int inner_maxit, int max_outer,
                               int outer_restart, vp_gmres_info* infos) {
  VP_API_BEGIN
  VP_REQUIRE(outer_restart > 0, ERR_INVALID, "outer_restart must be positive");
  solve_batched(ops, precs, nsolve, side, b, x, tol, restart, inner_tol, inner_maxit, max_outer, outer_restart,
                infos);
  VP_API_END
}

}  // extern "C"
