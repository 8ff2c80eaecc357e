// Per-kernel timing with HIP events on the launching stream.  Enabled per
// context; every instrumented launch records an event pair, resolved when
// statistics are read.
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include "op.hpp"

namespace vp {

namespace {
struct Pending {
  int kid;
  hipEvent_t a, b;
};
struct ProfState {
  bool on = false;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  long long launches[kNumKernelIds] = {0};
  double ms[kNumKernelIds] = {0};
};
std::mutex g_mu;
std::map<hipStream_t, ProfState> g_prof;

hipEvent_t take(ProfState& st) {
  if (!st.pool.empty()) {
    hipEvent_t e = st.pool.back();
    st.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void resolve(ProfState& st) {
  for (const Pending& p : st.pending) {
    float t = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
      st.launches[p.kid] += 1;
      st.ms[p.kid] += t;
    }
    st.pool.push_back(p.a);
    st.pool.push_back(p.b);
  }
  st.pending.clear();
}
}  // namespace

const char* kernel_id_name(int kid) {
  static const char* names[kNumKernelIds] = {"xfwd", "plane", "xinv", "pc_xfwd", "pc_plane", "pc_xinv",
                                                 "yfwd", "zpass", "yinv", "pc_yrows", "pc_block"};
  return (kid >= 0 && kid < kNumKernelIds) ? names[kid] : "";
}

void prof_begin(hipStream_t s, int kid, void** token) {
  *token = nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_prof.find(s);
  if (it == g_prof.end() || !it->second.on) return;
  ProfState& st = it->second;
  hipEvent_t a = take(st), b = take(st);
  if (!a || !b) return;
  (void)hipEventRecord(a, s);
  st.pending.push_back({kid, a, b});
  *token = reinterpret_cast<void*>(st.pending.size());   // 1-based index
}

void prof_end(hipStream_t s, void* token) {
  if (!token) return;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_prof.find(s);
  if (it == g_prof.end()) return;
  ProfState& st = it->second;
  const size_t i = reinterpret_cast<size_t>(token) - 1;
  if (i < st.pending.size()) (void)hipEventRecord(st.pending[i].b, s);
  if (st.pending.size() > 4096) {
    (void)hipStreamSynchronize(s);
    resolve(st);
  }
}

void prof_enable(hipStream_t s, bool on) {
  std::lock_guard<std::mutex> lk(g_mu);
  ProfState& st = g_prof[s];
  if (!on) {
    (void)hipStreamSynchronize(s);
    resolve(st);
  }
  st.on = on;
}

bool prof_enabled_on(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_prof.find(s);
  return it != g_prof.end() && it->second.on;
}

void prof_stats(hipStream_t s, int kid, long long* launches, double* ms, bool reset) {
  (void)hipStreamSynchronize(s);
  std::lock_guard<std::mutex> lk(g_mu);
  ProfState& st = g_prof[s];
  resolve(st);
  *launches = st.launches[kid];
  *ms = st.ms[kid];
  if (reset) {
    st.launches[kid] = 0;
    st.ms[kid] = 0;
  }
}

}  // namespace vp
