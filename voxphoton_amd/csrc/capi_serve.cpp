// C ABI: request server with continuous batching of mixed-precision solves.
//
// Requests wait in a FIFO queue and are admitted while the HBM budget (their
// solver state) and the slot count allow; an oversize request is admitted
// alone.  Admitted requests are grouped by operator geometry and solve
// parameters (a group key).  A few executor threads, each with its own HIP
// stream (VP_SERVER_STREAMS, default 3), take turns on the keys: an executor
// takes the key of the smallest requests, aged by the rounds it has waited
// (or the key that has waited longest), that no other executor is running,
// admits what the queue allows, and runs one refinement step (ir_step_set:
// FP64 residual, inner FP32 GMRES, update) for every active request of that
// key as one batched launch sequence.  Each request keeps the recycled Krylov
// space of its first inner solve between rounds (VP_SERVER_RECYCLE=0: none).
// Groups of different keys therefore
// run at the same time on the device; requests admitted later join their
// group at its next step, and finished requests leave and free their budget
// at once.  Per-problem arithmetic in the batched kernels never depends on
// the other problems of a launch, so a request's result is the same alone
// and in any mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "krylov.hpp"
#include "capi_common.hpp"

using namespace vp;


namespace {
using Clock = std::chrono::steady_clock;

struct ServeReq {
  int64_t ticket = 0;
  vp_op_s* op = nullptr;
  vp_prec_s* prec = nullptr;
  const double2* b = nullptr;
  double2* x = nullptr;
  vp_solve_params prm{};
  int64_t bytes = 0;
  std::string key;
  IrProblem st;
  bool finished = false;
  int error = 0;
  std::string errmsg;
  Clock::time_point t_submit, t_done;
  int64_t admitted_round = 0;
  // recycled Krylov space of the request's first inner solve (C and U,
  // right-aligned in kcap = restart columns of complex64), kept between rounds
  DevBuf rC, rU;
  int rkc = 0;
};

// VP_SERVER_RECYCLE=0: every inner solve of a served request starts afresh.
bool server_recycle() {
  const char* e = std::getenv("VP_SERVER_RECYCLE");
  return !(e && e[0] == '0');
}

// Solver-state bytes of one request when it runs in a batch: Krylov basis
// and GMRES vectors (FP32), refinement and gathered vectors (FP64), and the
// per-problem matvec / preconditioner workspaces of a batched launch.
int64_t request_bytes(const vp_op_s* op, const vp_prec_s* pc, const vp_solve_params& prm) {
  const int64_t n = op->n;
  int64_t b = int64_t(prm.restart + 1 + 6) * n * 8 + int64_t(6) * n * 16;
  // recycled space: its columns ahead of the basis in a round, and C and U kept between rounds
  if (server_recycle()) b += int64_t(3) * prm.restart * n * 8;
  const Geometry& g = op->p32.present ? op->p32.g : op->p64.g;
  const int64_t w = int64_t(3) * g.x.P * op->Mb * op->Nb;
  const int64_t w2 = op->p32.plan.variant == 1 ? int64_t(3) * g.x.P * g.y.P * op->Nb : 0;
  b += (w + w2) * (16 + 8);
  for (const vp_prec_s* q : prec_leaves(const_cast<vp_prec_s*>(pc)))
    b += int64_t(3) * q->g.V * (16 + 8) * (q->g.level == 1 ? 2 : 1);
  return b;
}

std::string group_key(const vp_op_s* o, const vp_prec_s* q, const vp_solve_params& p) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%d %d %d %d %d %d|%d %d %d %d %d %d|%lld %d %d %d %d %.17g|", o->p64.g.x.P,
                o->p64.g.y.P, o->p64.g.z.P, o->p32.g.x.P, o->p32.g.y.P, o->p32.g.z.P, o->l0,
                o->m0, o->n0, o->Lb, o->Mb, o->Nb, static_cast<long long>(o->n), o->structured,
                o->p64.plan.variant, o->p32.plan.variant, o->p32.plan.fsym, o->scale32);
  std::string k(buf);
  if (!q) k += "P-|";
  for (const vp_prec_s* l : prec_leaves(const_cast<vp_prec_s*>(q))) {
    std::snprintf(buf, sizeof(buf), "P%d %d %d %lld %d %lld %d %016llx %016llx|", l->g.L, l->g.M, l->g.N,
                  static_cast<long long>(l->nbox), l->g.level, static_cast<long long>(l->g.n_outer), l->g.bsym,
                  static_cast<unsigned long long>(l->pos_hash), static_cast<unsigned long long>(l->sub_hash));
    k += buf;
  }
  std::snprintf(buf, sizeof(buf), "%.17g %d %.17g %d %d %d", p.tol, p.restart, p.inner_tol, p.inner_maxit,
                p.max_outer, p.side);
  return k + buf;
}
}  // namespace

struct vp_server_s {
  int device = 0;
  vp_ctx ctx = nullptr;                 // the creating context (its recycle width applies to rounds)
  std::vector<hipStream_t> streams;     // one per executor
  int64_t budget = 0;
  int max_active = 1;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ServeReq*> pending;
  std::vector<ServeReq*> active;
  std::map<int64_t, std::unique_ptr<ServeReq>> reqs;
  std::map<std::string, int64_t> busy;  // key -> executor step number while it runs
  std::map<std::string, int64_t> last;  // key -> step number of its last start
  int64_t used = 0, next_ticket = 1, rounds = 0;
  bool stop = false;
  bool small_first = true;              // false: VP_SERVER_POLICY=rr
  std::vector<std::thread> workers;
};

namespace {

// One refinement step for a group of admitted requests of one key, on stream st.
void run_round(const std::vector<ServeReq*>& g, hipStream_t st, int recycle_width) {
  const int np = int(g.size());
  vp_op_s* o0 = g[0]->op;
  const int64_t n = o0->n;
  const vp_solve_params& prm = g[0]->prm;
  std::vector<vp_op_s*> ops(np);
  std::vector<vp_prec_s*> precs(np);
  for (int p = 0; p < np; ++p) {
    ops[p] = g[p]->op;
    precs[p] = g[p]->prec;
  }
  // workspaces of their own: another executor may run the same operator
  ProblemSet<double> p64;
  make_problem_set<double>(p64, ops.data(), nullptr, np, true);
  p64.stream = st;
  ProblemSet<float> p32;
  make_problem_set<float>(p32, ops.data(), precs[0] ? precs.data() : nullptr, np, true);
  p32.stream = st;
  p32.recycle_width = recycle_width;
  DevBuf B(size_t(n) * np * sizeof(double2)), X(size_t(n) * np * sizeof(double2));
  for (int p = 0; p < np; ++p) {
    VP_HIP(hipMemcpyAsync(B.as<double2>() + p * n, g[p]->b, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
    if (g[p]->st.steps == 0 && g[p]->st.hist.empty())
      VP_HIP(hipMemsetAsync(X.as<double2>() + p * n, 0, n * sizeof(double2), st));
    else
      VP_HIP(hipMemcpyAsync(X.as<double2>() + p * n, g[p]->x, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
  }
  std::vector<IrProblem> stt(np);
  for (int p = 0; p < np; ++p) stt[p] = g[p]->st;
  // each request's recycled space enters the round's problem set (zero where
  // a request has none yet: that request builds its own in this round) and
  // leaves it again, so its inner solves depend on its own history only
  Recycle rec;
  const bool recycle = server_recycle();
  const int kcap = prm.restart;
  const int64_t vprob = int64_t(2 * kcap + 1) * n, colb = int64_t(kcap) * n * int64_t(sizeof(float2));
  if (recycle) {
    rec.alloc(n, np, kcap, sizeof(float2));
    rec.kcp.assign(np, 0);
    rec.want.assign(np, 0);
    for (int p = 0; p < np; ++p) {
      float2* cdst = rec.V.as<float2>() + p * vprob;
      float2* udst = rec.U.as<float2>() + p * int64_t(kcap) * n;
      if (g[p]->rkc > 0) {
        VP_HIP(hipMemcpyAsync(cdst, g[p]->rC.p, colb, hipMemcpyDeviceToDevice, st));
        VP_HIP(hipMemcpyAsync(udst, g[p]->rU.p, colb, hipMemcpyDeviceToDevice, st));
      } else {
        VP_HIP(hipMemsetAsync(cdst, 0, colb, st));
        VP_HIP(hipMemsetAsync(udst, 0, colb, st));
      }
      rec.kcp[p] = g[p]->rkc;
      rec.want[p] = g[p]->rkc == 0 ? 1 : 0;
      rec.kc = std::max(rec.kc, g[p]->rkc);
    }
  }
  ir_step_set(p64, p32, B.as<double2>(), X.as<double2>(), o0->scale32, prm.tol, prm.restart, prm.inner_tol,
              prm.inner_maxit, prm.max_outer, prm.side, stt, recycle ? &rec : nullptr);
  for (int p = 0; p < np; ++p) {
    VP_HIP(hipMemcpyAsync(g[p]->x, X.as<double2>() + p * n, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
    if (recycle && rec.kcp[p] > 0 && g[p]->rkc == 0 && !stt[p].done) {
      g[p]->rC.alloc(size_t(colb));
      g[p]->rU.alloc(size_t(colb));
      VP_HIP(hipMemcpyAsync(g[p]->rC.p, rec.V.as<float2>() + p * vprob, colb, hipMemcpyDeviceToDevice, st));
      VP_HIP(hipMemcpyAsync(g[p]->rU.p, rec.U.as<float2>() + p * int64_t(kcap) * n, colb, hipMemcpyDeviceToDevice,
                            st));
      g[p]->rkc = rec.kcp[p];
    }
  }
  VP_HIP(hipStreamSynchronize(st));
  for (int p = 0; p < np; ++p)
    if (stt[p].done) {   // a finished request's space is freed at once
      g[p]->rC.release();
      g[p]->rU.release();
      g[p]->rkc = 0;
    }
  // the new states last, once the round's device work has finished: a
  // request seen as done may be collected (and deleted) by its client
  for (int p = 0; p < np; ++p) g[p]->st = stt[p];
}

// Admission (S->mu held): FIFO while the budget and the slots allow; a
// request is admitted alone when nothing else is active.
void admit(vp_server_s* S) {
  while (!S->pending.empty() && int(S->active.size()) < S->max_active) {
    ServeReq* r = S->pending.front();
    if (!S->active.empty() && S->used + r->bytes > S->budget) break;
    S->pending.pop_front();
    r->admitted_round = S->rounds;
    S->used += r->bytes;
    S->active.push_back(r);
  }
}

// The next key to run (S->mu held); "" if none runnable.  Default: the key
// with the fewest unknowns per request, its size halved for every 16 rounds
// it has waited (shortest work first, with ageing so that no key starves).
// VP_SERVER_POLICY=rr: the key that started longest ago (round robin).
std::string pick_key(vp_server_s* S) {
  std::string best;
  double best_score = 0.0;
  for (ServeReq* r : S->active) {
    // a busy key's requests belong to the executor running their round
    if (S->busy.count(r->key) || r->st.done) continue;
    auto it = S->last.find(r->key);
    const int64_t l = it == S->last.end() ? -1 : it->second;
    double score = double(l);
    if (S->small_first) {
      const double waited = double(S->rounds - std::max<int64_t>(l, r->admitted_round));
      score = double(r->op->n) * std::exp2(-waited / 16.0);
    }
    if (best.empty() || score < best_score) {
      best = r->key;
      best_score = score;
    }
  }
  return best;
}

void executor(vp_server_s* S, int k) {
  (void)hipSetDevice(S->device);
  hipStream_t st = S->streams[size_t(k)];
  for (;;) {
    std::vector<ServeReq*> batch;
    std::string key;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        admit(S);
        key = pick_key(S);
        if (!key.empty()) break;
        if (S->stop && S->pending.empty() && S->active.empty()) return;
        S->cv_work.wait(lk);
      }
      for (ServeReq* r : S->active)
        if (r->key == key && !r->st.done) batch.push_back(r);
      S->rounds += 1;
      S->busy[key] = S->rounds;
      S->last[key] = S->rounds;
    }
    try {
      run_round(batch, st, S->ctx->recycle_width);
    } catch (const VpError& e) {
      (void)hipStreamSynchronize(st);   // the failed round's queued work ends before its requests do
      for (ServeReq* r : batch) {
        r->error = e.code;
        r->errmsg = e.msg;
        r->st.done = true;
      }
    } catch (const std::exception& e) {
      (void)hipStreamSynchronize(st);
      for (ServeReq* r : batch) {
        r->error = VP_ERR_DEVICE;
        r->errmsg = e.what();
        r->st.done = true;
      }
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->busy.erase(key);
      const auto now = Clock::now();
      // only this round's requests: another executor may be running a round
      // of other requests, whose states are its own until that round ends
      std::vector<ServeReq*> keep;
      bool any_done = false;
      for (ServeReq* r : S->active) {
        if (std::find(batch.begin(), batch.end(), r) != batch.end() && r->st.done) {
          r->finished = true;
          r->t_done = now;
          S->used -= r->bytes;
          any_done = true;
        } else {
          keep.push_back(r);
        }
      }
      S->active.swap(keep);
      bool key_left = false;
      for (ServeReq* r : S->active) key_left |= r->key == key;
      if (!key_left) S->last.erase(key);
      (void)any_done;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();   // freed budget, a released key or a finished run
  }
}

void fill_serve_info(const ServeReq* r, vp_gmres_info* info) {
  if (!info) return;
  info->flag = r->error ? 1 : r->st.flag;
  info->relres = r->st.relres;
  info->ir_steps = r->st.steps;
  info->total_iters = r->st.inner;
  info->iter_inner = 0;
  info->iter_outer = r->st.steps;
  info->seconds = std::chrono::duration<double>(r->t_done - r->t_submit).count();
  info->resvec_len = 0;
  if (info->resvec && info->resvec_cap > 0) {
    const int cnt = std::min<int>(info->resvec_cap, int(r->st.hist.size()));
    for (int i = 0; i < cnt; ++i) info->resvec[i] = r->st.hist[i];
    info->resvec_len = cnt;
  }
}

ServeReq* find_req(vp_server_s* S, int64_t ticket) {
  auto it = S->reqs.find(ticket);
  VP_REQUIRE(it != S->reqs.end(), ERR_INVALID, "unknown server ticket");
  return it->second.get();
}
}  // namespace

extern "C" {

int vp_server_create(vp_ctx ctx, int64_t hbm_budget_bytes, int max_active, vp_server* out) {
  VP_API_BEGIN
  VP_REQUIRE(ctx && out && hbm_budget_bytes > 0 && max_active > 0, ERR_INVALID, "bad server arguments");
  VP_HIP(hipSetDevice(ctx->device));
  auto S = std::make_unique<vp_server_s>();
  S->device = ctx->device;
  S->ctx = ctx;
  S->budget = hbm_budget_bytes;
  S->max_active = max_active;
  int ns = 3;
  if (const char* e = std::getenv("VP_SERVER_STREAMS")) ns = std::max(1, std::min(8, std::atoi(e)));
  if (const char* e = std::getenv("VP_SERVER_POLICY")) S->small_first = std::strcmp(e, "rr") != 0;
  for (int k = 0; k < ns; ++k) {
    hipStream_t st = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) {
      for (hipStream_t o : S->streams) {
        unregister_stream(o);
        (void)hipStreamDestroy(o);
      }
      VP_HIP(e);
    }
    register_stream(S->device, st);
    S->streams.push_back(st);
  }
  vp_server_s* raw = S.get();
  try {
    for (int k = 0; k < ns; ++k) S->workers.emplace_back(executor, raw, k);
  } catch (...) {
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& w : S->workers)
      if (w.joinable()) w.join();
    for (hipStream_t o : S->streams) {
      unregister_stream(o);
      (void)hipStreamDestroy(o);
    }
    throw;
  }
  *out = S.release();
  VP_API_END
}

int vp_server_submit(vp_server s, vp_op op, vp_prec M, const void* b, void* x, const vp_solve_params* prm,
                     int64_t* ticket) {
  VP_API_BEGIN
  VP_REQUIRE(s && op && b && x && prm && ticket, ERR_INVALID, "bad submit arguments");
  VP_REQUIRE(prm->restart > 0 && prm->restart <= 1000 && prm->inner_maxit > 0 && prm->max_outer >= 0 &&
                 prm->tol > 0 && prm->inner_tol > 0 && (prm->side == VP_PREC_LEFT || prm->side == VP_PREC_RIGHT),
             ERR_INVALID, "bad solve parameters");
  VP_REQUIRE(op->p64.present && op->p32.present, ERR_INVALID,
             "mixed precision needs operators created with VP_FP64|VP_FP32");
  VP_REQUIRE(op->ctx->device == s->device, ERR_INVALID, "operator lives on another device");
  if (M) {
    const int64_t outer = M->g.n_outer ? M->g.n_outer : M->nbox;
    VP_REQUIRE(outer == op->n, ERR_INVALID, "preconditioner size does not match the operator");
  }
  auto r = std::make_unique<ServeReq>();
  r->op = op;
  r->prec = M;
  r->b = static_cast<const double2*>(b);
  r->x = static_cast<double2*>(x);
  r->prm = *prm;
  r->bytes = request_bytes(op, M, *prm);
  r->key = group_key(op, M, *prm);
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    VP_REQUIRE(!s->stop, ERR_INVALID, "server is shutting down");
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    s->pending.push_back(r.get());
    s->reqs.emplace(r->ticket, std::move(r));
  }
  s->cv_work.notify_one();
  VP_API_END
}

int vp_server_poll(vp_server s, int64_t ticket, int* done, vp_gmres_info* info) {
  VP_API_BEGIN
  VP_REQUIRE(s && done, ERR_INVALID, "bad poll arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  ServeReq* r = find_req(s, ticket);
  *done = r->finished ? 1 : 0;
  if (r->finished) {
    fill_serve_info(r, info);
    VP_REQUIRE(r->error == 0, r->error, "request failed: " + r->errmsg);
  }
  VP_API_END
}

int vp_server_wait(vp_server s, int64_t ticket, vp_gmres_info* info) {
  VP_API_BEGIN
  VP_REQUIRE(s, ERR_INVALID, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  ServeReq* r = find_req(s, ticket);
  s->cv_done.wait(lk, [&] { return r->finished; });
  fill_serve_info(r, info);
  const int err = r->error;
  const std::string msg = r->errmsg;
  s->reqs.erase(ticket);
  VP_REQUIRE(err == 0, err, "request failed: " + msg);
  VP_API_END
}

int vp_server_stats(vp_server s, int64_t* pending, int64_t* active, int64_t* bytes_in_use, int64_t* rounds) {
  VP_API_BEGIN
  VP_REQUIRE(s, ERR_INVALID, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  if (pending) *pending = int64_t(s->pending.size());
  if (active) *active = int64_t(s->active.size());
  if (bytes_in_use) *bytes_in_use = s->used;
  if (rounds) *rounds = s->rounds;
  VP_API_END
}

int vp_server_destroy(vp_server s) {
  VP_API_BEGIN
  if (!s) return VP_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& w : s->workers)
    if (w.joinable()) w.join();
  (void)hipSetDevice(s->device);
  for (hipStream_t st : s->streams) {
    (void)hipStreamSynchronize(st);
    unregister_stream(st);
    (void)hipStreamDestroy(st);
  }
  delete s;
  VP_API_END
}

}  // extern "C"
