// Workgroup-level mixed-radix FFTs staged through LDS (gfx950).
//
// A batch of 1D DFTs lives in LDS; every stage of the Stockham autosort
// algorithm (Govindaraju et al., "High performance discrete Fourier transforms
// on graphics processors", SC'08) reads each butterfly's R inputs from one
// LDS buffer, twiddles, runs a radix-R DFT in registers and writes the other
// buffer.
// Radices 2, 3, 4, 5, 7, 8, 11, 13 and 16 are supported.  Forward transforms
// use exp(-2*pi*i*jk/n), inverse ones exp(+2*pi*i*jk/n); neither normalises.
#pragma once
#include "common.hpp"

namespace vp {

// ---------------------------------------------------------------------------
// Compile-time sines and cosines of 2*pi*num/den: octant reduction to
// [0, pi/4] followed by Taylor series (exact to a few ulp in double).
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) { term *= -x * x / double((2 * k) * (2 * k + 1)); sum += term; }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) { term *= -x * x / double((2 * k - 1) * (2 * k)); sum += term; }
  return sum;
}
struct SinCos { double s, c; };
constexpr SinCos sincos_turn(long num, long den) {
  long n = ((num % den) + den) % den;
  long q = 8 * n;
  long o = q / den;             // octant 0..7
  long r = q - o * den;         // remainder in [0, den)
  double sa = 0, ca = 0;
  if ((o & 1) == 0) {
    double phi = (kPi / 4.0) * double(r) / double(den);
    sa = taylor_sin(phi); ca = taylor_cos(phi);
  } else {
    double psi = (kPi / 4.0) * double(den - r) / double(den);
    sa = taylor_cos(psi); ca = taylor_sin(psi);
  }
  long Q = o / 2;
  SinCos out{0, 0};
  if (Q == 0) { out.s = sa; out.c = ca; }
  else if (Q == 1) { out.s = ca; out.c = -sa; }
  else if (Q == 2) { out.s = -sa; out.c = -ca; }
  else { out.s = -ca; out.c = sa; }
  return out;
}

template <int R> struct RootTable {
  double c[R], s[R];
  constexpr RootTable() : c(), s() {
    for (int i = 0; i < R; ++i) { SinCos v = sincos_turn(i, R); c[i] = v.c; s[i] = v.s; }
  }
};

// ---------------------------------------------------------------------------
// Radix-R butterflies on registers (in place).
// ---------------------------------------------------------------------------
template <bool INV, typename V>
__device__ __forceinline__ void bf2(V& a, V& b) { V t = a; a = cadd(t, b); b = csub(t, b); }

template <bool INV, typename V>
__device__ __forceinline__ void dft4(V* v) {
  V s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  V s13 = cadd(v[1], v[3]), d13 = mul_mi<INV>(csub(v[1], v[3]));
  v[0] = cadd(s02, s13); v[2] = csub(s02, s13);
  v[1] = cadd(d02, d13); v[3] = csub(d02, d13);
}

template <typename T, bool INV>
__device__ __forceinline__ C<T> rot8(C<T> a, int k) {
  // a * exp(-+ 2 pi i k / 8), k in 0..3
  const T h = T(0.70710678118654752440084436210484903928);
  if (k == 0) return a;
  if (k == 2) return mul_mi<INV>(a);
  C<T> w;
  if (k == 1) { w.x = h; w.y = INV ? h : -h; }
  else { w.x = -h; w.y = INV ? h : -h; }
  return cmul(a, w);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(C<T>* v) {
  C<T> e[4] = {v[0], v[2], v[4], v[6]};
  C<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e); dft4<INV>(o);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    C<T> t = rot8<T, INV>(o[k], k);
    v[k] = cadd(e[k], t); v[k + 4] = csub(e[k], t);
  }
}

template <typename T, bool INV>
__device__ __forceinline__ void dft16(C<T>* v) {
  C<T> e[8], o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) { e[k] = v[2 * k]; o[k] = v[2 * k + 1]; }
  dft8<T, INV>(e); dft8<T, INV>(o);
  constexpr RootTable<16> rt{};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    C<T> w; w.x = T(rt.c[k]); w.y = INV ? T(rt.s[k]) : T(-rt.s[k]);
    C<T> t = cmul(o[k], w);
    v[k] = cadd(e[k], t); v[k + 8] = csub(e[k], t);
  }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_odd(C<T>* v) {
  constexpr int H = (R - 1) / 2;
  constexpr RootTable<R> rt{};
  C<T> a[H], b[H];
#pragma unroll
  for (int j = 1; j <= H; ++j) { a[j - 1] = cadd(v[j], v[R - j]); b[j - 1] = csub(v[j], v[R - j]); }
  C<T> x0 = v[0];
  C<T> y0 = x0;
#pragma unroll
  for (int j = 0; j < H; ++j) y0 = cadd(y0, a[j]);
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    T re = x0.x, im = x0.y, sr = T(0), si = T(0);
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const T c = T(rt.c[(j * k) % R]);
      const T s = T(rt.s[(j * k) % R]);
      re += a[j - 1].x * c; im += a[j - 1].y * c;
      sr += s * b[j - 1].y; si -= s * b[j - 1].x;
    }
    if (INV) { sr = -sr; si = -si; }
    v[k] = mk<T>(re + sr, im + si);
    v[R - k] = mk<T>(re - sr, im - si);
  }
  v[0] = y0;
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(C<T>* v) {
  if constexpr (R == 2) bf2<INV>(v[0], v[1]);
  else if constexpr (R == 4) dft4<INV>(v);
  else if constexpr (R == 8) dft8<T, INV>(v);
  else if constexpr (R == 16) dft16<T, INV>(v);
  else dft_odd<T, R, INV>(v);
}

// Pruned radix-R DFTs for zero-padded transforms, rc uniform over the
// workgroup (1 <= rc <= R):
//   dft_in_pruned : only inputs v[r], r < rc, are nonzero; all R outputs.
//   dft_out_pruned: all R inputs; only outputs q < rc are wanted (odd R by
//                   the symmetric pairs of dft_odd; the rest are left as is).
template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_in_pruned(C<T>* v, int rc) {
  constexpr RootTable<R> rt{};
  C<T> o[R];
#pragma unroll
  for (int k = 0; k < R; ++k) o[k] = v[0];
#pragma unroll
  for (int r = 1; r < R; ++r) {
    if (r < rc) {
      o[0] = cadd(o[0], v[r]);
#pragma unroll
      for (int k = 1; k < R; ++k) {
        const int e = (r * k) % R;
        o[k] = cadd(o[k], cmul(v[r], mk<T>(T(rt.c[e]), INV ? T(rt.s[e]) : T(-rt.s[e]))));
      }
    }
  }
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = o[k];
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_out_pruned(C<T>* v, int rc) {
  if constexpr (R % 2 == 0) {
    dft<T, R, INV>(v);
  } else {
    constexpr int H = (R - 1) / 2;
    constexpr RootTable<R> rt{};
    C<T> a[H], b[H];
#pragma unroll
    for (int j = 1; j <= H; ++j) { a[j - 1] = cadd(v[j], v[R - j]); b[j - 1] = csub(v[j], v[R - j]); }
    const C<T> x0 = v[0];
    C<T> y0 = x0;
#pragma unroll
    for (int j = 0; j < H; ++j) y0 = cadd(y0, a[j]);
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      if (k < rc || R - k < rc) {
        T re = x0.x, im = x0.y, sr = T(0), si = T(0);
#pragma unroll
        for (int j = 1; j <= H; ++j) {
          const T c = T(rt.c[(j * k) % R]);
          const T sn = T(rt.s[(j * k) % R]);
          re += a[j - 1].x * c; im += a[j - 1].y * c;
          sr += sn * b[j - 1].y; si -= sn * b[j - 1].x;
        }
        if (INV) { sr = -sr; si = -si; }
        v[k] = mk<T>(re + sr, im + si);
        v[R - k] = mk<T>(re - sr, im - si);
      }
    }
    v[0] = y0;
  }
}

// ---------------------------------------------------------------------------
// Division by a runtime-invariant divisor d (1 <= d < 2^31) for numerators
// n < 2^31: q = (umulhi(n, m) + n) >> s with s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1 (Granlund & Montgomery, PLDI'94).
// ---------------------------------------------------------------------------
struct FastDiv {
  uint32_t d, m, s;
  __host__ __device__ FastDiv() : d(1), m(1), s(0) {}
  __host__ __device__ explicit FastDiv(uint32_t dd) { init(dd); }
  __host__ __device__ void init(uint32_t dd) {
    d = dd;
    s = 0;
    while ((uint64_t(1) << s) < d) ++s;
    m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
  }
  __device__ __forceinline__ uint32_t div(uint32_t n) const {
    return uint32_t((uint64_t(__umulhi(n, m)) + n) >> s);
  }
  __device__ __forceinline__ uint32_t mod(uint32_t n) const { return n - div(n) * d; }
};

// Cheap division for index arithmetic inside kernels: numerators 0 <= n < 2^20.
// n * rcp(d) in FP32 is within 1/8 of n/d there, so one correction step
// gives the exact quotient.
struct IDiv {
  int d;
  float inv;
  __device__ __forceinline__ explicit IDiv(int dd) : d(dd), inv(__builtin_amdgcn_rcpf(float(dd))) {}
  __device__ __forceinline__ int div(int n) const {
    int q = int(float(n) * inv);
    const int r = n - q * d;
    q += (r >= d) - (r < 0);
    return q;
  }
  __device__ __forceinline__ int mod(int n) const { return n - div(n) * d; }
};

// ---------------------------------------------------------------------------
// Plans and LDS layouts
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 12;

// One 1D DFT length.  The twiddle table passed alongside holds
// exp(-2*pi*i*k/P) for k < P with P = n * tws.  The radix of stage s sits in
// bits [5s, 5s+5) of ``code`` (a scalar, so kernels never index an array).
struct FftPlan {
  int n;
  int nst;
  int tws;
  int pad;
  unsigned long long code;
  __host__ __device__ __forceinline__ int radix(int s) const { return int((code >> (5 * s)) & 31ull); }
};

// Batch layout in LDS: transform beta starts at (beta / bi) * so + (beta % bi) * si
// and its element j sits at start + j * st.
struct LdsLayout {
  int bi, so, si, st;
  __device__ __forceinline__ int off(int beta) const {
    int hi = beta / bi;
    return hi * so + (beta - hi * bi) * si;
  }
};

template <typename T, bool INV>
__device__ __forceinline__ C<T> twiddle(const C<T>* __restrict__ tw, int idx) {
  C<T> w = tw[idx];
  if (INV) w.y = -w.y;
  return w;
}

// Twiddle w^idx, w = exp(-2 pi i / n), from two short FP64 factor tables in
// LDS, lo[a] = w^a (a < 64) and hi[b] = w^(64 b): w^idx = lo[idx % 64] *
// hi[idx / 64], formed in FP64 and rounded once to T, so FP32 twiddles equal
// the correctly rounded table entries.  fill() evaluates the (64 +
// ceil(n/64)) entries with sincospi; every thread of the workgroup calls it
// and a barrier must follow.
template <typename T>
struct TwSplit {
  const double2* lo;
  const double2* hi;
  __device__ __forceinline__ C<T> operator()(int idx) const {
    const double2 a = lo[idx & 63], b = hi[idx >> 6];
    C<T> w;
    w.x = T(a.x * b.x - a.y * b.y);
    w.y = T(a.x * b.y + a.y * b.x);
    return w;
  }
  static constexpr int entries(int n) { return 64 + (n + 63) / 64; }   // double2 entries
  __device__ static TwSplit fill(void* lds, int n) {
    double2* t = reinterpret_cast<double2*>(lds);
    const int nhi = (n + 63) / 64;
    for (int e = threadIdx.x; e < 64 + nhi; e += blockDim.x) {
      const int k = e < 64 ? e : 64 * (e - 64);
      double sv, cv;
      sincospi(2.0 * double(k) / double(n), &sv, &cv);
      t[e] = make_double2(cv, -sv);
    }
    return TwSplit{t, t + 64};
  }
};

template <typename T, bool INV>
__device__ __forceinline__ C<T> twiddle(const TwSplit<T>& tw, int idx) {
  C<T> w = tw(idx);
  if (INV) w.y = -w.y;
  return w;
}

// One Stockham stage from ``src`` to ``dst`` (distinct LDS buffers with the
// same layout).  Each thread loops over butterflies, so only R values are
// live at a time; the caller synchronises between stages.
template <typename T, int R, bool INV>
__device__ __forceinline__ void stage(const C<T>* src, C<T>* dst, int n, int Ns, int nbatch, const LdsLayout& lay,
                                      const C<T>* __restrict__ tw, int twmul) {
  const int nb = n / R;
  const int total = nb * nbatch;
  const IDiv dnb(nb), dns(Ns), dbi(lay.bi);
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    const int beta = dnb.div(b);
    const int j = b - beta * nb;
    const int hi = dbi.div(beta);
    const int off = hi * lay.so + (beta - hi * lay.bi) * lay.si;
    C<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[off + (j + r * nb) * lay.st];
    const int jm = dns.mod(j);
    if (Ns > 1) {
      // w = exp(-+ 2 pi i jm / (Ns R)) = table[jm * (n / (Ns R)) * tws]
      const int step = jm * twmul;
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<T, INV>(tw, r * step));
    }
    dft<T, R, INV>(v);
    const int base = (j - jm) * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[off + (base + r * Ns) * lay.st] = v[r];
  }
}

// Batch-fastest variant: element j of transform beta sits at beta + j * ld
// and consecutive threads take consecutive transforms, so every LDS access
// of a wavefront (same j, consecutive beta) is a contiguous run of words.
template <typename T, int R, bool INV>
__device__ __forceinline__ void stage_bf(const C<T>* src, C<T>* dst, int n, int Ns, int nbatch, int ld,
                                         const C<T>* __restrict__ tw, int twmul) {
  const int nb = n / R;
  const int total = nb * nbatch;
  const IDiv dnbatch(nbatch), dns(Ns);
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    const int j = dnbatch.div(b);
    const int beta = b - j * nbatch;
    C<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[beta + (j + r * nb) * ld];
    const int jm = dns.mod(j);
    if (Ns > 1) {
      const int step = jm * twmul;
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<T, INV>(tw, r * step));
    }
    dft<T, R, INV>(v);
    const int base = (j - jm) * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[beta + (base + r * Ns) * ld] = v[r];
  }
}

// Pruned stage of a zero-padded transform (batch-fastest layout), with a
// compile-time count RC:
//   first stage (Ns = 1): inputs j + r*nb with r >= RC are zero padding
//     (the caller guarantees it), so every output sums RC terms;
//   last stage (Ns * R = n): only outputs j + r*Ns with r < RC are wanted.
template <typename T, int R, int RC, bool INV, bool FIRST>
__device__ __forceinline__ void stage_bf_pruned(const C<T>* src, C<T>* dst, int n, int Ns, int nbatch, int ld,
                                                const C<T>* __restrict__ tw, int twmul) {
  constexpr RootTable<R> rt{};
  const int nb = n / R;
  const int total = nb * nbatch;
  const IDiv dnbatch(nbatch), dns(Ns);
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    const int j = dnbatch.div(b);
    const int beta = b - j * nbatch;
    if (FIRST) {
      C<T> v[RC];
#pragma unroll
      for (int r = 0; r < RC; ++r) v[r] = src[beta + (j + r * nb) * ld];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        C<T> acc = v[0];
#pragma unroll
        for (int r = 1; r < RC; ++r) {
          const int e = (r * k) % R;
          C<T> w;
          w.x = T(rt.c[e]);
          w.y = INV ? T(rt.s[e]) : T(-rt.s[e]);
          acc = cadd(acc, cmul(v[r], w));
        }
        dst[beta + (j * R + k) * ld] = acc;   // Ns = 1: outputs j*R + k
      }
    } else {
      C<T> v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = src[beta + (j + r * nb) * ld];
      const int jm = dns.mod(j);
      if (Ns > 1) {
        const int step = jm * twmul;
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<T, INV>(tw, r * step));
      }
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < RC; ++r) {
        C<T> acc = v[0];
#pragma unroll
        for (int k = 1; k < R; ++k) {
          const int e = (r * k) % R;
          C<T> w;
          w.x = T(rt.c[e]);
          w.y = INV ? T(rt.s[e]) : T(-rt.s[e]);
          acc = cadd(acc, cmul(v[k], w));
        }
        dst[beta + (base + r * Ns) * ld] = acc;
      }
    }
  }
}

// Runs one stage, pruned when mode != 0 and (R, rc) has an instantiation:
// mode 1 first stage with rc nonzero inputs per butterfly, mode 2 last stage
// with rc wanted outputs per butterfly.
template <typename T, bool INV>
__device__ __forceinline__ void stage_bf_any(const C<T>* src, C<T>* dst, int R, int n, int Ns, int nbatch, int ld,
                                             const C<T>* __restrict__ tw, int twmul, int mode, int rc) {
#define VP_PR(RR, K)                                                                                    \
  if (R == RR && rc == K) {                                                                             \
    if (mode == 1) stage_bf_pruned<T, RR, K, INV, true>(src, dst, n, Ns, nbatch, ld, tw, twmul);          \
    else stage_bf_pruned<T, RR, K, INV, false>(src, dst, n, Ns, nbatch, ld, tw, twmul);                  \
    return;                                                                                             \
  }
  if (mode != 0) {
    VP_PR(11, 1) VP_PR(11, 2) VP_PR(11, 3) VP_PR(11, 4) VP_PR(11, 5)
    VP_PR(7, 1) VP_PR(7, 2) VP_PR(7, 3)
    VP_PR(5, 1) VP_PR(5, 2)
    VP_PR(8, 1) VP_PR(8, 2) VP_PR(8, 3) VP_PR(8, 4)
  }
#undef VP_PR
  switch (R) {
    case 2: stage_bf<T, 2, INV>(src, dst, n, Ns, nbatch, ld, tw, twmul); break;
    case 3: stage_bf<T, 3, INV>(src, dst, n, Ns, nbatch, ld, tw, twmul); break;
    case 4: stage_bf<T, 4, INV>(src, dst, n, Ns, nbatch, ld, tw, twmul); break;
    case 5: stage_bf<T, 5, INV>(src, dst, n, Ns, nbatch, ld, tw, twmul); break;
    case 7: stage_bf<T, 7, INV>(src, dst, n, Ns, nbatch, ld, tw, twmul); break;
    case 8: stage_bf<T, 8, INV>(src, dst, n, Ns, nbatch, ld, tw, twmul); break;
    case 11: stage_bf<T, 11, INV>(src, dst, n, Ns, nbatch, ld, tw, twmul); break;
    case 13:
      if constexpr (sizeof(T) == 4) stage_bf<T, 13, INV>(src, dst, n, Ns, nbatch, ld, tw, twmul);
      break;
    case 16:
      if constexpr (sizeof(T) == 4) stage_bf<T, 16, INV>(src, dst, n, Ns, nbatch, ld, tw, twmul);
      break;
    default: break;
  }
}

// nin: leading inputs that may be nonzero (the rest are zero padding);
// nout: leading outputs needed.  Values >= n disable the pruning.
template <typename T, bool INV>
__device__ C<T>* lds_fft_bf(C<T>* a, C<T>* b, const FftPlan& pl, int nbatch, int ld, const C<T>* __restrict__ tw,
                            int nin = 1 << 30, int nout = 1 << 30) {
  int Ns = 1;
  C<T>* src = a;
  C<T>* dst = b;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.radix(s);
    const int twmul = (pl.n / (Ns * R)) * pl.tws;
    // pruned counts are uniform over the stage: ceil(lim / span) terms (first
    // stage, span nb = n/R) or outputs (last stage, span Ns)
    int mode = 0, rc = 0;
    if (s == 0 && nin < pl.n && pl.nst > 1) { mode = 1; const int nb = pl.n / R; rc = (nin + nb - 1) / nb; }
    else if (s == pl.nst - 1 && nout < pl.n && pl.nst > 1) { mode = 2; rc = (nout + Ns - 1) / Ns; }
    if (rc >= R) mode = 0;
    stage_bf_any<T, INV>(src, dst, R, pl.n, Ns, nbatch, ld, tw, twmul, mode, rc);
    __syncthreads();
    C<T>* t = src; src = dst; dst = t;
    Ns *= R;
  }
  return src;
}

// Runs all stages of ``pl`` on ``nbatch`` transforms laid out per ``lay``,
// ping-ponging between ``a`` and ``b``; returns the buffer holding the
// result.  Ends with a barrier.  Radices 13 and 16 are single precision only.
template <typename T, bool INV>
__device__ C<T>* lds_fft(C<T>* a, C<T>* b, const FftPlan& pl, int nbatch, const LdsLayout& lay,
                         const C<T>* __restrict__ tw) {
  int Ns = 1;
  C<T>* src = a;
  C<T>* dst = b;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.radix(s);
    const int twmul = (pl.n / (Ns * R)) * pl.tws;
    switch (R) {
      case 2: stage<T, 2, INV>(src, dst, pl.n, Ns, nbatch, lay, tw, twmul); break;
      case 3: stage<T, 3, INV>(src, dst, pl.n, Ns, nbatch, lay, tw, twmul); break;
      case 4: stage<T, 4, INV>(src, dst, pl.n, Ns, nbatch, lay, tw, twmul); break;
      case 5: stage<T, 5, INV>(src, dst, pl.n, Ns, nbatch, lay, tw, twmul); break;
      case 7: stage<T, 7, INV>(src, dst, pl.n, Ns, nbatch, lay, tw, twmul); break;
      case 8: stage<T, 8, INV>(src, dst, pl.n, Ns, nbatch, lay, tw, twmul); break;
      case 11: stage<T, 11, INV>(src, dst, pl.n, Ns, nbatch, lay, tw, twmul); break;
      case 13:
        if constexpr (sizeof(T) == 4) stage<T, 13, INV>(src, dst, pl.n, Ns, nbatch, lay, tw, twmul);
        break;
      case 16:
        if constexpr (sizeof(T) == 4) stage<T, 16, INV>(src, dst, pl.n, Ns, nbatch, lay, tw, twmul);
        break;
      default: break;
    }
    __syncthreads();
    C<T>* t = src; src = dst; dst = t;
    Ns *= R;
  }
  return src;
}

// ---------------------------------------------------------------------------
// Compile-time Stockham FFT over a batch of NB transforms of length NLEN in
// LDS, element k of transform b at k*LD + b.  Every size is a template
// constant, so index arithmetic folds into shifts and multiplies and the
// butterfly loops unroll.  A stage maps thread it to butterfly j of
// transform b either j-fastest (it = b*(NLEN/R) + j: a wavefront walks
// consecutive elements of one transform) or b-fastest (it = j*NB + b:
// consecutive transforms at one element); LD odd keeps both free of LDS
// bank conflicts for 8-byte elements.  The first stage reads through
// ld(b, k) and the last one writes through st(b, k, v), so global data can
// enter and leave in whichever order is contiguous in memory (FB / LB pick
// b-fastest there).  tw[k * twm] = exp(-2 pi i k / NLEN).
// ---------------------------------------------------------------------------
template <typename T, bool INV, int NLEN, int NB, int LD, int NT, int NS, int R, bool FIRST, bool LAST, bool BF,
          class Tw, class Ld, class St>
__device__ __forceinline__ void ct_stage(const C<T>* src, C<T>* dst, const Tw& tw, int twm, Ld& ld, St& st) {
  constexpr int NBF = NLEN / R;
  constexpr int TOTAL = NBF * NB;
#pragma unroll
  for (int it0 = 0; it0 < TOTAL; it0 += NT) {
    const int it = it0 + int(threadIdx.x);
    if (TOTAL % NT == 0 || it < TOTAL) {
      const int b = BF ? it % NB : it / NBF;
      const int j = BF ? it / NB : it - b * NBF;
      C<T> v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = FIRST ? ld(b, j + NBF * r) : src[(j + NBF * r) * LD + b];
      const int jm = NS > 1 ? j % NS : 0;
      if constexpr (NS > 1) {
        constexpr int TM = NLEN / (NS * R);
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<T, INV>(tw, r * jm * TM * twm));
      }
      dft<T, R, INV>(v);
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (LAST) st(b, base + r * NS, v[r]);
        else dst[(base + r * NS) * LD + b] = v[r];
      }
    }
  }
}

// Runs the radix list R0, Rs... (product NLEN) from ``ld`` to ``st`` (or,
// LDSOUT, to LDS), ping-ponging between buffers a and b with a barrier
// between stages.  Returns the LDS buffer holding the result when LDSOUT
// (after a barrier), else nullptr.
template <typename T, bool INV, int NLEN, int NB, int LD, int NT, int NS, bool FIRST, bool FB, bool LB, bool LDSOUT,
          int R0, int... Rs, class Tw, class Ld, class St>
__device__ __forceinline__ C<T>* ct_fft(C<T>* a, C<T>* b, const Tw& tw, int twm, Ld& ld, St& st) {
  constexpr bool END = sizeof...(Rs) == 0;
  constexpr bool BF = FIRST ? FB : (END && !LDSOUT ? LB : false);
  ct_stage<T, INV, NLEN, NB, LD, NT, NS, R0, FIRST, END && !LDSOUT, BF>(a, b, tw, twm, ld, st);
  if constexpr (!END) {
    __syncthreads();
    return ct_fft<T, INV, NLEN, NB, LD, NT, NS * R0, false, FB, LB, LDSOUT, Rs...>(b, a, tw, twm, ld, st);
  } else if constexpr (LDSOUT) {
    __syncthreads();
    return b;
  } else {
    return nullptr;
  }
}

// In-place form of ct_stage / ct_fft: one LDS buffer.  A stage first loads
// every butterfly input of the thread (ITER x R values in registers), then,
// after a barrier, writes its outputs into the same buffer, so a transform
// needs half the LDS of the ping-pong form (the first stage reads through
// ld and needs no barrier before its writes).
template <typename T, bool INV, int NLEN, int NB, int LD, int NT, int NS, int R, bool FIRST, bool LAST, bool BF,
          class Tw, class Ld, class St>
__device__ __forceinline__ void ct_stage_ip(C<T>* buf, const Tw& tw, int twm, Ld& ld, St& st) {
  constexpr int NBF = NLEN / R;
  constexpr int TOTAL = NBF * NB;
  constexpr int ITER = (TOTAL + NT - 1) / NT;
  C<T> v[ITER][R];
#pragma unroll
  for (int i = 0; i < ITER; ++i) {
    const int it = i * NT + int(threadIdx.x);
    if (TOTAL % NT == 0 || it < TOTAL) {
      const int b = BF ? it % NB : it / NBF;
      const int j = BF ? it / NB : it - b * NBF;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = FIRST ? ld(b, j + NBF * r) : buf[(j + NBF * r) * LD + b];
    }
  }
  if constexpr (!FIRST) __syncthreads();
#pragma unroll
  for (int i = 0; i < ITER; ++i) {
    const int it = i * NT + int(threadIdx.x);
    if (TOTAL % NT == 0 || it < TOTAL) {
      const int b = BF ? it % NB : it / NBF;
      const int j = BF ? it / NB : it - b * NBF;
      const int jm = NS > 1 ? j % NS : 0;
      if constexpr (NS > 1) {
        constexpr int TM = NLEN / (NS * R);
#pragma unroll
        for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], twiddle<T, INV>(tw, r * jm * TM * twm));
      }
      dft<T, R, INV>(v[i]);
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (LAST) st(b, base + r * NS, v[i][r]);
        else buf[(base + r * NS) * LD + b] = v[i][r];
      }
    }
  }
}

template <typename T, bool INV, int NLEN, int NB, int LD, int NT, int NS, bool FIRST, bool FB, bool LB, bool LDSOUT,
          int R0, int... Rs, class Tw, class Ld, class St>
__device__ __forceinline__ C<T>* ct_fft_ip(C<T>* buf, const Tw& tw, int twm, Ld& ld, St& st) {
  constexpr bool END = sizeof...(Rs) == 0;
  constexpr bool BF = FIRST ? FB : (END && !LDSOUT ? LB : false);
  ct_stage_ip<T, INV, NLEN, NB, LD, NT, NS, R0, FIRST, END && !LDSOUT, BF>(buf, tw, twm, ld, st);
  if constexpr (!END) {
    __syncthreads();
    return ct_fft_ip<T, INV, NLEN, NB, LD, NT, NS * R0, false, FB, LB, LDSOUT, Rs...>(buf, tw, twm, ld, st);
  } else if constexpr (LDSOUT) {
    __syncthreads();
    return buf;
  } else {
    return nullptr;
  }
}

}  // namespace vp
