// Host-side planning: radix factorisation and launch shapes for the matvec.
#include <algorithm>
#include <map>
#include <cstdlib>
#include <mutex>
#include <vector>
#include "vp_internal.hpp"

namespace vp {

bool factorize(int n, bool fp64, FftPlan& out, int tws, bool odd_first) {
  out.n = n;
  out.tws = tws;
  out.nst = 0;
  out.pad = 0;
  out.code = 0;
  int m = n;
  auto push = [&](int r) {
    if (out.nst >= kMaxStages) return false;
    out.code |= (unsigned long long)r << (5 * out.nst);
    out.nst++;
    m /= r;
    return true;
  };
  const int odd32[] = {13, 11, 7, 5, 3};
  const int odd64[] = {11, 7, 5, 3};
  const int* odd = fp64 ? odd64 : odd32;
  const int nodd = fp64 ? 4 : 5;
  auto odds = [&]() {
    for (int i = 0; i < nodd; ++i)
      while (m % odd[i] == 0) if (!push(odd[i])) return false;
    return true;
  };
  // largest odd radices first when asked (a zero-padded input then prunes
  // the most expensive stage), else powers of two first
  if (odd_first && !odds()) return false;
  // powers of two: 16 (single precision only), 8, 4, then a final 2
  if (!fp64) while (m % 16 == 0) if (!push(16)) return false;
  while (m % 8 == 0) if (!push(8)) return false;
  while (m % 4 == 0) if (!push(4)) return false;
  while (m % 2 == 0) if (!push(2)) return false;
  if (!odd_first && !odds()) return false;
  return m == 1;
}

int max_lds_bytes(int device) {
  static std::mutex mu;
  static int cache[64] = {0};
  std::lock_guard<std::mutex> lk(mu);
  if (device < 0 || device >= 64) return 65536;
  if (!cache[device]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0)
      v = 65536;
    cache[device] = v;
  }
  return cache[device];
}

static int pow2ceil(int v) {
  int p = 1;
  while (p < v) p *= 2;
  return p;
}

// Radix pairs (RA, RB) of the fused z kernel, instantiated in matvec.hip in
// the same order; the first pair with RA * RB = Np is used.
static const int kZfPairs[][3] = {   // RA, RB, single precision only
    {11, 8, 0}, {11, 4, 0}, {11, 2, 0}, {7, 8, 0}, {7, 4, 0}, {7, 2, 0}, {5, 8, 0}, {5, 4, 0}, {5, 2, 0},
    {3, 8, 0},  {3, 4, 0},  {3, 2, 0},  {3, 3, 0}, {8, 8, 0}, {8, 4, 0}, {4, 4, 0}, {4, 2, 0},
    {11, 16, 1}, {7, 16, 1}, {5, 16, 1}, {3, 16, 1}, {16, 8, 1}};

bool zfused_pair(int Np, bool fp64, int& ra, int& rb) {
  for (const auto& p : kZfPairs) {
    if (p[0] * p[1] != Np || (fp64 && p[2])) continue;
    ra = p[0];
    rb = p[1];
    return true;
  }
  return false;
}

bool yct_supported(int n) {
#define VP_YCT_CASE(NL, ...) case NL:
  switch (n) {
    VP_YCT_LIST(VP_YCT_CASE)
    return true;
    default: return false;
  }
#undef VP_YCT_CASE
}

bool pcrow_supported(int n) {
#define VP_PCROW_CASE(NL, ...) case NL:
  switch (n) {
    VP_PCROW_LIST(VP_PCROW_CASE)
    return true;
    default: return false;
  }
#undef VP_PCROW_CASE
}

namespace {
constexpr size_t kCacheMin = size_t(64) << 20;
constexpr size_t kCacheMax = size_t(48) << 30;
// A released buffer is handed out again once the events recorded at its
// release on every stream of the library (contexts, servers) on its device
// have completed: all work enqueued before the release, which is all work
// that can use it, has then finished.  No device-wide synchronisation.
struct CachedBuf {
  int dev;
  void* p;
  size_t bytes;
  std::vector<hipEvent_t> ev;
};
std::mutex g_cache_mu;
std::vector<CachedBuf> g_cache;               // released, oldest first
std::map<void*, std::pair<int, size_t>> g_big;  // live and cached large allocations: device, size
size_t g_cache_bytes = 0;
std::vector<std::pair<int, hipStream_t>> g_streams;   // the library's streams (under g_cache_mu)

bool events_done(const CachedBuf& c) {
  for (hipEvent_t e : c.ev)
    if (hipEventQuery(e) != hipSuccess) return false;
  return true;
}

void drop(CachedBuf& c) {   // g_cache_mu held
  for (hipEvent_t e : c.ev) {
    (void)hipEventSynchronize(e);
    (void)hipEventDestroy(e);
  }
  c.ev.clear();
  (void)hipFree(c.p);
  g_big.erase(c.p);
  g_cache_bytes -= c.bytes;
}

void flush_device(int dev) {   // g_cache_mu held
  for (size_t i = 0; i < g_cache.size();) {
    if (dev < 0 || g_cache[i].dev == dev) {
      drop(g_cache[i]);
      g_cache.erase(g_cache.begin() + long(i));
    } else {
      ++i;
    }
  }
}
}  // namespace

void register_stream(int device, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_streams.emplace_back(device, s);
}

void unregister_stream(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  for (size_t i = 0; i < g_streams.size(); ++i)
    if (g_streams[i].second == s) {
      g_streams.erase(g_streams.begin() + long(i));
      break;
    }
}

void* dev_alloc(size_t b) {
  int dev = 0;
  VP_HIP(hipGetDevice(&dev));
  if (b >= kCacheMin) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    int best = -1;
    for (int i = 0; i < int(g_cache.size()); ++i) {
      const CachedBuf& c = g_cache[size_t(i)];
      if (c.dev == dev && c.bytes >= b && c.bytes - c.bytes / 5 <= b &&
          (best < 0 || c.bytes < g_cache[size_t(best)].bytes) && events_done(c))
        best = i;
    }
    if (best >= 0) {
      CachedBuf c = g_cache[size_t(best)];
      for (hipEvent_t e : c.ev) (void)hipEventDestroy(e);
      g_cache_bytes -= c.bytes;
      g_cache.erase(g_cache.begin() + best);
      return c.p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, b);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lk(g_cache_mu);
      flush_device(dev);
    }
    e = hipMalloc(&p, b);
  }
  VP_HIP(e);
  if (b >= kCacheMin) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_big[p] = {dev, b};
  }
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_big.find(p);
  if (it == g_big.end()) {
    (void)hipFree(p);
    return;
  }
  CachedBuf c{it->second.first, p, it->second.second, {}};
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (cur != c.dev) (void)hipSetDevice(c.dev);
  bool ok = true;
  for (const auto& ds : g_streams) {
    if (ds.first != c.dev) continue;
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, ds.second) != hipSuccess) {
      if (ev) (void)hipEventDestroy(ev);
      ok = false;
      break;
    }
    c.ev.push_back(ev);
  }
  if (!ok) {   // no events: fall back to waiting for the device
    for (hipEvent_t ev : c.ev) (void)hipEventDestroy(ev);
    c.ev.clear();
    (void)hipDeviceSynchronize();
  }
  if (cur != c.dev) (void)hipSetDevice(cur);
  g_cache.push_back(c);
  g_cache_bytes += c.bytes;
  while (g_cache_bytes > kCacheMax && !g_cache.empty()) {
    drop(g_cache.front());
    g_cache.erase(g_cache.begin());
  }
}

void dev_cache_flush() {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  flush_device(-1);
}

int xct_threads(int S, bool fp64) {
#define VP_XCT_CASE(SL, NTF, FR, NTI, IR) \
  case SL: return NTF;
  if (fp64) {
    switch (S) {
      VP_XCT64_LIST(VP_XCT_CASE)
      default: return 0;
    }
  }
  switch (S) {
    VP_XCT32_LIST(VP_XCT_CASE)
    default: return 0;
  }
#undef VP_XCT_CASE
}

// Mirrors acc_bound<T, EPT>() in matvec.hip.
static int acc_bound(bool fp64, int ept) { return (fp64 ? ept <= 1 : ept <= 4) ? 1024 : 512; }

// Smallest accumulator count whose thread count fits its launch bound.
static bool pick_threads(int points, bool fp64, int& t_out, int& e_out) {
  const int ept_max = 32;
  for (int e = 1; e <= ept_max; e *= 2) {
    int t = (points + e - 1) / e;
    t = std::max(64, ((t + 63) / 64) * 64);
    if (t <= acc_bound(fp64, e)) { t_out = t; e_out = e; return true; }
  }
  return false;
}

bool make_matvec_plan(const Geometry& g, bool fp64, int device, MatvecPlan& out, std::string& why, int variant) {
  const size_t s = fp64 ? 16 : 8;
  const int lds_max = max_lds_bytes(device);
  out.variant = variant;
  if (variant == 1) {
    // ---- split passes: y sweeps over By pencils, z sweep over Tz ky columns
    const int Mp = g.y.P, Np = g.z.P;
    // LDS per workgroup: about 40 KB (four workgroups per CU) measured best
    // for the z pass in both precisions; VP_SPLIT_LDS overrides (bytes)
    size_t budget = 40 * 1024;
    if (const char* e = std::getenv("VP_SPLIT_LDS")) budget = size_t(std::max(4096, std::atoi(e)));
    out.By = 0;
    for (int B : {16, 8, 4, 2, 1}) {
      const int ld = Mp + (B > 1 ? 1 : 0);
      const size_t lds = 2 * size_t(B) * ld * s;
      if (lds <= budget || B == 1) { out.By = B; out.ldy = ld; out.lds_y = lds; break; }
    }
    out.Tz = 0;
    int tz_cap = 32;
    if (const char* e = std::getenv("VP_ZTILE")) tz_cap = std::max(1, std::atoi(e));
    for (int Tz : {32, 16, 8, 4, 2, 1}) {
      if (Tz > Mp || Tz > tz_cap) continue;
      const size_t lds = (2 * size_t(3) * Tz * Np + Np) * s;
      if (lds <= budget || Tz == 1) { out.Tz = Tz; out.lds_z = lds; break; }
    }
    if (out.lds_y > size_t(lds_max) || out.lds_z > size_t(lds_max)) { why = "split passes exceed LDS"; return false; }
    // long lines (one per workgroup): more threads per stage
    out.ty = (out.By == 1 && Mp >= 1024) ? 512 : 256;
    out.tz = 256;
    // z transforms of the split pass: forward with the odd radices first
    // (pruned zero-padded input), inverse with them last (pruned outputs)
    out.zi = g.z.plan;
    if (!factorize(Np, fp64, out.zf, 1, true)) { why = "z FFT length does not factor"; return false; }
    // register z-column kernel (VP_ZCOL=1): Np = lanes x in-register radix; the LDS pass is the default
    out.zr1 = out.zr2 = 0;
    const char* ez = std::getenv("VP_ZCOL");
    if (ez && ez[0] == '1') {
      for (int r1 : {8, 4, 16}) {
        if (Np % r1) continue;
        const int r2 = Np / r1;
        if (r2 == 3 || r2 == 4 || r2 == 5 || r2 == 7 || r2 == 8 || r2 == 11) {
          out.zr1 = r1;
          out.zr2 = r2;
          break;
        }
      }
    }
    // compile-time y pass where Mp has one (VP_YCT=0 keeps the generic pass)
    const char* ey = std::getenv("VP_YCT");
    out.yct = (!(ey && ey[0] == '0') && yct_supported(Mp) && yct_lds(Mp, int(s)) <= lds_max) ? Mp : 0;
    // compile-time x passes over pencil-major x spectra where the class
    // length has them and the y pass is compile-time too (VP_XCT=0 keeps
    // the generic x passes and plane-major x spectra)
    const char* ex = std::getenv("VP_XCT");
    out.xct = 0;
    // LDS: one class buffer (the stages run in place) plus the FP64 split
    // twiddle tables (64 + ceil(Lp/64) entries)
    const size_t xlds = size_t(g.x.S) * s + (64 + (size_t(g.x.P) + 63) / 64) * 16;
    if (out.yct && !(ex && ex[0] == '0') && xct_threads(g.x.S, fp64) > 0 && xlds <= size_t(lds_max))
      out.xct = g.x.S;
    // fused two-stage z kernel (default; VP_ZFUSED=0 keeps the LDS z pass)
    out.zra = out.zrb = out.ztpk = 0;
    out.lds_zf = 0;
    const char* ef = std::getenv("VP_ZFUSED");
    if (out.zr1 == 0 && !(ef && ef[0] == '0')) {
      int ra = 0, rb = 0;
      if (zfused_pair(Np, fp64, ra, rb)) {
        const size_t lds = size_t(Np) * 3 * kZfTile * s;
        if (lds <= size_t(lds_max)) {
          out.zra = ra;
          out.zrb = rb;
          out.ztpk = zfused_threads(ra, rb);
          out.lds_zf = lds;
        }
      }
    }
    out.t2 = 0;
    out.ept2 = 1;
    out.lds2 = 0;
  }
  // ---- K2 (planes): classes fixed by the caller; pick threads/accumulators.
  if (variant == 0) {
    const int Ms = g.y.S, Ns = g.z.S;
    const long long E = 3LL * Ms * Ns;
    const long long lds = (2 * E + g.y.P + g.z.P) * (long long)s;   // work buffers + twiddles
    long long cap = lds_max;
    if (const char* e = std::getenv("VP_K2_LDS_CAP")) cap = std::min<long long>(cap, std::atoll(e));
    if (lds > cap) { why = "plane class buffers exceed LDS"; return false; }
    int t = 0, e = 0;
    if (!pick_threads(3 * g.M * g.N, fp64, t, e)) { why = "plane too large for register accumulators"; return false; }
    out.t2 = t;
    out.ept2 = e;
    out.lds2 = size_t(lds);
  }
  // ---- K1/K3 (x pencils): largest pencil batch that fits LDS and registers.
  {
    const int Ls = g.x.S;
    bool ok = false;
    // at most two pencils per workgroup: more workgroups in flight measured
    // faster than wider ones (VP_XPENCILS overrides)
    int bcap = 2;
    if (const char* e = std::getenv("VP_XPENCILS")) bcap = std::max(1, std::atoi(e));
    for (int B : {16, 8, 4, 2, 1}) {
      if (B > 1 && (B > g.M || B > bcap)) continue;
      const int ld = Ls + (B > 1 ? 1 : 0);
      const size_t lds = 2 * size_t(B) * ld * s;
      if (lds > size_t(lds_max)) continue;
      int t = 0, e = 0;
      if (!pick_threads(B * g.L, fp64, t, e)) continue;
      if (e > 16) continue;
      out.B = B; out.ldsx = ld; out.tx = std::max(t, 256); out.ept3 = e; out.lds1 = lds;
      if (out.tx > acc_bound(fp64, e)) out.tx = acc_bound(fp64, e);
      ok = true;
      break;
    }
    if (!ok) { why = "no pencil tiling fits the x FFT"; return false; }
  }
  return true;
}

}  // namespace vp
