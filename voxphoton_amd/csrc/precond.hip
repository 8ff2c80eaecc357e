// Block-circulant (Chan) preconditioners on the device.
//
// Apply (pc/src_preconditioning/chan_2_mvp_for_parallel_idx.m:2-23):
//   scatter -> DFT along x (length L) -> per kx: DFT along y (length M),
//   3N x 3N block product per (kx, ky), inverse DFT along y -> inverse DFT
//   along x -> gather.  The 1/(L*M) of MATLAB's ifft pair is folded into
//   the stored blocks.
// 1-level apply (chan_mvp_idx.m:2-16) is the same with M folded into the
// block (3MN x 3MN blocks, no y transform).
// Setup (level_2_parallel_func.m:1-192 semantics) builds every 3N x 3N
// block directly from the Toeplitz tensor: Chan circulant + DFT along x,
// signed y-Toeplitz first row/column per (c, n_t, d, n_s), Chan circulant +
// DFT along y, then batched Gauss-Jordan inversion in FP64.
#include <dlfcn.h>
#include <rocsolver/rocsolver.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>
#include "precond.hpp"

namespace vp {

// ----------------------------------------------------------------------------
// Apply kernels
// ----------------------------------------------------------------------------
// acc += (br + i bi) * x as four FMAs
template <typename T>
__device__ __forceinline__ void cfma_acc(C<T>& acc, T br, T bi, C<T> x) {
  acc.x = fma(br, x.x, acc.x);
  acc.x = fma(-bi, x.y, acc.x);
  acc.y = fma(br, x.y, acc.y);
  acc.y = fma(bi, x.x, acc.y);
}

template <typename T>
__device__ __forceinline__ int64_t pc_outer(const PcArgs<T>& a, int64_t j) {
  return a.sub ? int64_t(a.sub[j]) : j;
}

// x pass forward: B pencils (consecutive m) at fixed (n, c), length L DFT.
template <typename T>
__global__ void __launch_bounds__(512) k_pc_xfwd(PcArgs<T> a, PcGeom g) {
  if (a.active && !a.active[blockIdx.y]) return;
  const C<T>* __restrict__ X = a.x + int64_t(blockIdx.y) * a.xstride;
  C<T>* __restrict__ work = a.work + int64_t(blockIdx.y) * a.wstride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);
  const int L = g.L, M = g.M, N = g.N, B = g.B, ld = g.ld;
  C<T>* W1 = W0 + B * ld;
  const int nmb = (M + B - 1) / B;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int mb = blk % nmb, rest = blk / nmb;
  const int n = rest % N, c = rest / N;
  const int m0 = mb * B;
  const int64_t MN = int64_t(M) * N;
  for (int e = threadIdx.x; e < B * L; e += blockDim.x) {
    const int bb = e / L, l = e - bb * L, m = m0 + bb;
    C<T> v = czero<T>();
    if (m < M) {
      const int64_t vox = l + int64_t(L) * (m + int64_t(M) * n);
      int64_t j;
      if (g.structured) {
        const int p = a.pos[vox];
        j = p < 0 ? -1 : p + int64_t(c) * g.nS;
      } else {
        j = a.pos[int64_t(c) * g.V + vox];
      }
      if (j >= 0) v = X[pc_outer(a, j)];
    }
    W0[bb * ld + l] = v;
  }
  __syncthreads();
  const C<T>* R = lds_fft<T, false>(W0, W1, g.px, B, LdsLayout{1, ld, 0, 1}, a.twx);
  for (int e = threadIdx.x; e < B * L; e += blockDim.x) {
    const int kx = e / B, bb = e - kx * B, m = m0 + bb;
    if (m < M) work[(int64_t(kx) * 3 + c) * MN + int64_t(n) * M + m] = R[bb * ld + kx];
  }
}

// per kx plane: y DFT, 3N x 3N block products, inverse y DFT.
template <typename T>
__global__ void __launch_bounds__(512) k_pc_plane(PcArgs<T> a, PcGeom g) {
  if (a.active && !a.active[blockIdx.y]) return;
  C<T>* __restrict__ work = a.work + int64_t(blockIdx.y) * a.wstride;
  const C<T>* __restrict__ blocks = a.blocks[blockIdx.y];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);
  const int M = g.M, N = g.N, n3 = 3 * N;
  const int plane = n3 * M;
  C<T>* W1 = W0 + plane;
  // launch order pairs kx with L - kx (0, 1, L-1, 2, L-2, ...), which share
  // their stored blocks when the blocks are kept as a quadrant
  const int L = g.L;
  const int io = xcd_remap(blockIdx.x, gridDim.x);
  const int jo = (io + 1) >> 1;
  const int kx = io == 0 ? 0 : ((io & 1) ? jo : L - jo);
  C<T>* Zp = work + int64_t(kx) * plane;
  for (int e = threadIdx.x; e < plane; e += blockDim.x) W0[e] = Zp[e];
  __syncthreads();
  const LdsLayout lay{1, M, 0, 1};   // transform q = (c, n): elements q*M + m
  C<T>* R = lds_fft<T, false>(W0, W1, g.py, n3, lay, a.twy);
  C<T>* O = (R == W0) ? W1 : W0;
  // One thread per (block row, ky): consecutive threads read consecutive rows
  // of a column-major block, so every load of the block stream is coalesced.
  // ky runs in mirror-pair order too, so B(kx, ky) and B(kx, M-ky) are read
  // by neighbouring wavefronts.  Mirrored blocks come from the stored one as
  // S B S: the x (y) component's row and column sums change sign.
  const bool mx = g.bsym && 2 * kx > L;
  const C<T>* Bk = blocks + int64_t(mx ? L - kx : kx) * g.Hy * n3 * n3;
  for (int e = threadIdx.x; e < plane; e += blockDim.x) {
    const int i = e / n3;
    const int r = e - i * n3;
    const int jy = (i + 1) >> 1;
    const int ky = i == 0 ? 0 : ((i & 1) ? jy : M - jy);
    const bool my = g.bsym && 2 * ky > M;
    const C<T>* Bb = Bk + int64_t(my ? M - ky : ky) * n3 * n3 + r;
    C<T> sx = czero<T>(), sy = czero<T>(), sz = czero<T>();
    for (int q = 0; q < N; ++q) {
      sx = cadd(sx, cmul(Bb[int64_t(q) * n3], R[q * M + ky]));
      sy = cadd(sy, cmul(Bb[int64_t(q + N) * n3], R[(q + N) * M + ky]));
      sz = cadd(sz, cmul(Bb[int64_t(q + 2 * N) * n3], R[(q + 2 * N) * M + ky]));
    }
    if (mx) sx = cneg(sx);
    if (my) sy = cneg(sy);
    C<T> acc = cadd(cadd(sx, sy), sz);
    const int cr = r / N;
    if ((cr == 0 && mx) || (cr == 1 && my)) acc = cneg(acc);
    O[r * M + ky] = acc;
  }
  __syncthreads();
  const C<T>* F = lds_fft<T, true>(O, R, g.py, n3, lay, a.twy);
  for (int e = threadIdx.x; e < plane; e += blockDim.x) Zp[e] = F[e];
}

// Split plane, pass 1/3: y DFT (length M) of Brow consecutive rows of the
// box spectrum (row (kx, c, n) holds M contiguous entries), in place.
template <typename T, bool INV>
__global__ void __launch_bounds__(256) k_pc_yrows(PcArgs<T> a, PcGeom g) {
  if (a.active && !a.active[blockIdx.y]) return;
  C<T>* __restrict__ work = a.work + int64_t(blockIdx.y) * a.wstride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);
  const int M = g.M, B = g.Brow, ld = g.ldrow;
  C<T>* W1 = W0 + B * ld;
  const int64_t nrows = int64_t(g.L) * 3 * g.N;
  const int64_t r0 = int64_t(xcd_remap(blockIdx.x, gridDim.x)) * B;
  const int nb = int(min<int64_t>(B, nrows - r0));
  const IDiv dM(M);
  for (int e = threadIdx.x; e < B * M; e += blockDim.x) {
    const int bb = dM.div(e), m = e - bb * M;
    W0[bb * ld + m] = bb < nb ? work[(r0 + bb) * M + m] : czero<T>();
  }
  __syncthreads();
  const C<T>* R = lds_fft<T, INV>(W0, W1, g.py, B, LdsLayout{1, ld, 0, 1}, a.twy);
  for (int e = threadIdx.x; e < nb * M; e += blockDim.x) {
    const int bb = dM.div(e), m = e - bb * M;
    work[(r0 + bb) * M + m] = R[bb * ld + m];
  }
}

// Split plane, pass 2/3: 3N x 3N block products for TK consecutive ky of one
// kx (kx paired with L - kx in launch order, as the stored quadrant of the
// blocks is shared).  The 3N x TK column tile is staged in LDS; thread
// (t, r) reads row r of block (kx, ky0 + t) column by column (coalesced).
template <typename T>
__global__ void __launch_bounds__(256) k_pc_blockmul(PcArgs<T> a, PcGeom g) {
  if (a.active && !a.active[blockIdx.y]) return;
  C<T>* __restrict__ work = a.work + int64_t(blockIdx.y) * a.wstride;
  const C<T>* __restrict__ blocks = a.blocks[blockIdx.y];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* X = reinterpret_cast<C<T>*>(smem_raw);   // X[q * TK + t]
  const int L = g.L, M = g.M, N = g.N, n3 = 3 * N, TK = g.TK;
  const int nt = (M + TK - 1) / TK;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int io = blk / nt;
  const int jo = (io + 1) >> 1;
  const int kx = io == 0 ? 0 : ((io & 1) ? jo : L - jo);
  const int ky0 = (blk - io * nt) * TK;
  C<T>* plane = work + int64_t(kx) * n3 * M;       // row q = c*N + n at plane + q*M
  const IDiv dTK(TK), dn3(n3);
  for (int e = threadIdx.x; e < n3 * TK; e += blockDim.x) {
    const int q = dTK.div(e), t = e - q * TK;
    X[e] = (ky0 + t < M) ? plane[int64_t(q) * M + ky0 + t] : czero<T>();
  }
  __syncthreads();
  const bool mx = g.bsym && 2 * kx > L;
  const C<T>* Bk = blocks + int64_t(mx ? L - kx : kx) * g.Hy * n3 * n3;
  for (int e = threadIdx.x; e < n3 * TK; e += blockDim.x) {
    const int t = dn3.div(e), r = e - t * n3;
    const int ky = ky0 + t;
    if (ky >= M) continue;
    const bool my = g.bsym && 2 * ky > M;
    const C<T>* Bb = Bk + int64_t(my ? M - ky : ky) * n3 * n3 + r;
    C<T> sx = czero<T>(), sy = czero<T>(), sz = czero<T>();
    for (int q = 0; q < N; ++q) {
      sx = cadd(sx, cmul(Bb[int64_t(q) * n3], X[q * TK + t]));
      sy = cadd(sy, cmul(Bb[int64_t(q + N) * n3], X[(q + N) * TK + t]));
      sz = cadd(sz, cmul(Bb[int64_t(q + 2 * N) * n3], X[(q + 2 * N) * TK + t]));
    }
    if (mx) sx = cneg(sx);
    if (my) sy = cneg(sy);
    C<T> acc = cadd(cadd(sx, sy), sz);
    const int cr = r / N;
    if ((cr == 0 && mx) || (cr == 1 && my)) acc = cneg(acc);
    plane[int64_t(r) * M + ky] = acc;
  }
}

// Index of the spectrum entry of row q = c*N + n at (kx, ky) in the layout
// the block products read: q-contiguous Y[kx][ky][q] (qlay, in work2),
// pencil-major X[q][ky][kx] (pct) or plane-major X[kx][q][ky].
__device__ __forceinline__ int64_t pc_spec_index(const PcGeom& g, int q, int ky, int kx, int n3) {
  if (g.qlay) return (int64_t(kx) * g.M + ky) * n3 + q;
  if (g.pct) return (int64_t(q) * g.M + ky) * g.L + kx;
  return (int64_t(kx) * n3 + q) * g.M + ky;
}

// Quadrant block products (blocks stored for kx <= L/2, ky <= M/2): one
// workgroup per stored block B(kxc, kyc) applies it to the up to four
// columns (kx, ky) in {kxc, L-kxc} x {kyc, M-kyc} that share it, so every
// block entry is read once and feeds four products.  The mirrored blocks
// are S B S (S negates the x or y component): the sign of the column's
// component is folded into the staged vector, the row sign into the result.
template <typename T, int QH = 8>
__global__ void __launch_bounds__(256) k_pc_quadmul(PcArgs<T> a, PcGeom g) {
  if (a.active && !a.active[blockIdx.y]) return;
  C<T>* __restrict__ work = (g.qlay ? a.work2 : a.work) + int64_t(blockIdx.y) * a.wstride;
  const C<T>* __restrict__ blocks = a.blocks[blockIdx.y];
  const int L = g.L, M = g.M, N = g.N, n3 = 3 * N, Hy = g.Hy;
  __shared__ C<T> Xs[4][3 * 85];
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  // consecutive workgroups (one XCD) take neighbouring columns: along ky for
  // plane-major spectra (ky contiguous), along kx for pencil-major (kx
  // contiguous), so their column gathers share cache lines in L2
  const int Hx = g.Hx;
  const int kxc = g.pct ? blk % Hx : blk / Hy;
  const int kyc = g.pct ? blk / Hx : blk - kxc * Hy;
  int kxs[4], kys[4];
  bool mxs[4], mys[4], use[4];
  const bool twox = kxc != 0 && 2 * kxc != L, twoy = kyc != 0 && 2 * kyc != M;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    mxs[c] = (c & 1) != 0;
    mys[c] = (c & 2) != 0;
    use[c] = (!mxs[c] || twox) && (!mys[c] || twoy);
    kxs[c] = mxs[c] ? L - kxc : kxc;
    kys[c] = mys[c] ? M - kyc : kyc;
  }
  for (int e = threadIdx.x; e < 4 * n3; e += blockDim.x) {
    const int c = e / n3, q = e - c * n3;
    C<T> v = czero<T>();
    if (use[c]) {
      v = work[pc_spec_index(g, q, kys[c], kxs[c], n3)];
      const int cq = q / N;
      if ((cq == 0 && mxs[c]) || (cq == 1 && mys[c])) v = cneg(v);
    }
    Xs[c][q] = v;
  }
  __syncthreads();
  const int64_t boff = (int64_t(kxc) * Hy + kyc) * n3 * n3;
  const C<T>* Bb = blocks + boff;
  const uint32_t* Hb = a.hblocks ? a.hblocks[blockIdx.y] + boff : nullptr;
  for (int r = threadIdx.x; r < n3; r += blockDim.x) {
    C<T> acc0 = czero<T>(), acc1 = czero<T>(), acc2 = czero<T>(), acc3 = czero<T>();
    if constexpr (sizeof(T) == 4) {
      if (Hb) {
        // binary16 blocks: half the bytes of the block stream.  Columns go in
        // chunks of QU whose block loads are all issued before any is used;
        // each entry is converted once and enters four complex
        // multiply-accumulates of two FMAs per part.
        auto step = [&](uint32_t hw, int q) {
          T br = T(__builtin_bit_cast(_Float16, uint16_t(hw & 0xffffu)));
          T bi = T(__builtin_bit_cast(_Float16, uint16_t(hw >> 16)));
          asm volatile("" : "+v"(br), "+v"(bi));   // keep the conversion out of the FMAs
          cfma_acc(acc0, br, bi, Xs[0][q]);
          cfma_acc(acc1, br, bi, Xs[1][q]);
          cfma_acc(acc2, br, bi, Xs[2][q]);
          cfma_acc(acc3, br, bi, Xs[3][q]);
        };
        constexpr int QU = QH;
        int q0 = 0;
        for (; q0 + QU <= n3; q0 += QU) {
          uint32_t hw[QU];
#pragma unroll
          for (int u = 0; u < QU; ++u) hw[u] = Hb[int64_t(q0 + u) * n3 + r];
#pragma unroll
          for (int u = 0; u < QU; ++u) step(hw[u], q0 + u);
        }
        for (; q0 < n3; ++q0) step(Hb[int64_t(q0) * n3 + r], q0);
        const T hs = T(a.hscale[blockIdx.y]);
        acc0 = cscale(acc0, hs);
        acc1 = cscale(acc1, hs);
        acc2 = cscale(acc2, hs);
        acc3 = cscale(acc3, hs);
      }
    }
    if (!Hb) {
      auto step = [&](C<T> b, int q) {
        cfma_acc(acc0, b.x, b.y, Xs[0][q]);
        cfma_acc(acc1, b.x, b.y, Xs[1][q]);
        cfma_acc(acc2, b.x, b.y, Xs[2][q]);
        cfma_acc(acc3, b.x, b.y, Xs[3][q]);
      };
      constexpr int QU = sizeof(T) == 8 ? 4 : 8;
      int q0 = 0;
      for (; q0 + QU <= n3; q0 += QU) {
        C<T> bv[QU];
#pragma unroll
        for (int u = 0; u < QU; ++u) bv[u] = Bb[int64_t(q0 + u) * n3 + r];
#pragma unroll
        for (int u = 0; u < QU; ++u) step(bv[u], q0 + u);
      }
      for (; q0 < n3; ++q0) step(Bb[int64_t(q0) * n3 + r], q0);
    }
    const int cr = r / N;
    const C<T> accs[4] = {acc0, acc1, acc2, acc3};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (!use[c]) continue;
      C<T> v = accs[c];
      if ((cr == 0 && mxs[c]) || (cr == 1 && mys[c])) v = cneg(v);
      work[pc_spec_index(g, r, kys[c], kxs[c], n3)] = v;
    }
  }
}

// FP32 quadrant block product with binary16 blocks on the packed dot
// instruction v_dot2_f32_f16 (two fp16 products summed into an f32
// accumulator): a stored 32-bit block word (re, im) enters
//   Re(b x) = dot2((b.re, b.im), (x.re, -x.im)),  Im(b x) = dot2((b.re, b.im), (x.im, x.re)),
// so an entry costs two instructions per column instead of a conversion and
// four FMAs.  The four mirrored columns are staged once per block as fp16
// pairs scaled per column by a power of two that puts max |x| at 2^14
// (exact scaling; fp16 keeps 11 significant bits, as the blocks do); the
// accumulators are rescaled by hscale / s_c.  Same layout, mirrors and sign
// flips as k_pc_quadmul.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256) k_pc_quadmul_dot2(PcArgs<float> a, PcGeom g) {
  if (a.active && !a.active[blockIdx.y]) return;
  float2* __restrict__ work = (g.qlay ? a.work2 : a.work) + int64_t(blockIdx.y) * a.wstride;
  const int L = g.L, M = g.M, N = g.N, n3 = 3 * N, Hy = g.Hy;
  // XX[q][2c], XX[q][2c+1]: the two fp16 pairs of column c at entry q
  __shared__ uint32_t XX[3 * 85][8];
  __shared__ float cmax[4][4];
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int Hx = g.Hx;
  const int kxc = g.pct ? blk % Hx : blk / Hy;
  const int kyc = g.pct ? blk / Hx : blk - kxc * Hy;
  int kxs[4], kys[4];
  bool mxs[4], mys[4], use[4];
  const bool twox = kxc != 0 && 2 * kxc != L, twoy = kyc != 0 && 2 * kyc != M;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    mxs[c] = (c & 1) != 0;
    mys[c] = (c & 2) != 0;
    use[c] = (!mxs[c] || twox) && (!mys[c] || twoy);
    kxs[c] = mxs[c] ? L - kxc : kxc;
    kys[c] = mys[c] ? M - kyc : kyc;
  }
  // gather the columns (at most 4 entries per thread: 4 * n3 <= 4 * 255 and
  // blockDim.x >= n3 rounded up to a wavefront) and their maxima
  float2 vals[4];
  float m4[4] = {0.f, 0.f, 0.f, 0.f};
  const int nel = 4 * n3;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = threadIdx.x + u * blockDim.x;
    float2 v = make_float2(0.f, 0.f);
    if (e < nel) {
      const int c = e / n3, q = e - c * n3;
      if (use[c]) {
        v = work[pc_spec_index(g, q, kys[c], kxs[c], n3)];
        const int cq = q / N;
        if ((cq == 0 && mxs[c]) || (cq == 1 && mys[c])) v = make_float2(-v.x, -v.y);
      }
      const float am = fmaxf(fabsf(v.x), fabsf(v.y));
#pragma unroll
      for (int c2 = 0; c2 < 4; ++c2)
        if (c2 == c) m4[c2] = fmaxf(m4[c2], am);
    }
    vals[u] = v;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    for (int off = 32; off > 0; off >>= 1) m4[c] = fmaxf(m4[c], __shfl_xor(m4[c], off));
    if (lane == 0) cmax[wave][c] = m4[c];
  }
  __syncthreads();
  float sc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float m = 0.f;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) m = fmaxf(m, cmax[w][c]);
    int ex = 0;
    (void)frexpf(m, &ex);                     // m < 2^ex
    sc[c] = m > 0.f ? ldexpf(1.f, 14 - ex) : 1.f;
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = threadIdx.x + u * blockDim.x;
    if (e < nel) {
      const int c = e / n3, q = e - c * n3;
      const float xr = vals[u].x * sc[c], xi = vals[u].y * sc[c];
      half2_t pa, pb;
      pa.x = _Float16(xr);
      pa.y = _Float16(-xi);
      pb.x = _Float16(xi);
      pb.y = _Float16(xr);
      XX[q][2 * c] = __builtin_bit_cast(uint32_t, pa);
      XX[q][2 * c + 1] = __builtin_bit_cast(uint32_t, pb);
    }
  }
  __syncthreads();
  const int64_t boff = (int64_t(kxc) * Hy + kyc) * n3 * n3;
  const uint32_t* Hb = a.hblocks[blockIdx.y] + boff;
  const float hs = a.hscale[blockIdx.y];
  for (int r = threadIdx.x; r < n3; r += blockDim.x) {
    float ar[4] = {0.f, 0.f, 0.f, 0.f}, ai[4] = {0.f, 0.f, 0.f, 0.f};
    auto step = [&](uint32_t hw, int q) {
      const half2_t b = __builtin_bit_cast(half2_t, hw);
      const uint4 x01 = *reinterpret_cast<const uint4*>(&XX[q][0]);
      const uint4 x23 = *reinterpret_cast<const uint4*>(&XX[q][4]);
      ar[0] = __builtin_amdgcn_fdot2(b, __builtin_bit_cast(half2_t, x01.x), ar[0], false);
      ai[0] = __builtin_amdgcn_fdot2(b, __builtin_bit_cast(half2_t, x01.y), ai[0], false);
      ar[1] = __builtin_amdgcn_fdot2(b, __builtin_bit_cast(half2_t, x01.z), ar[1], false);
      ai[1] = __builtin_amdgcn_fdot2(b, __builtin_bit_cast(half2_t, x01.w), ai[1], false);
      ar[2] = __builtin_amdgcn_fdot2(b, __builtin_bit_cast(half2_t, x23.x), ar[2], false);
      ai[2] = __builtin_amdgcn_fdot2(b, __builtin_bit_cast(half2_t, x23.y), ai[2], false);
      ar[3] = __builtin_amdgcn_fdot2(b, __builtin_bit_cast(half2_t, x23.z), ar[3], false);
      ai[3] = __builtin_amdgcn_fdot2(b, __builtin_bit_cast(half2_t, x23.w), ai[3], false);
    };
    constexpr int QU = 8;
    int q0 = 0;
    for (; q0 + QU <= n3; q0 += QU) {
      uint32_t hw[QU];
#pragma unroll
      for (int u = 0; u < QU; ++u) hw[u] = Hb[int64_t(q0 + u) * n3 + r];
#pragma unroll
      for (int u = 0; u < QU; ++u) step(hw[u], q0 + u);
    }
    for (; q0 < n3; ++q0) step(Hb[int64_t(q0) * n3 + r], q0);
    const int cr = r / N;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (!use[c]) continue;
      const float f = hs / sc[c];
      float2 v = make_float2(ar[c] * f, ai[c] * f);
      if ((cr == 0 && mxs[c]) || (cr == 1 && mys[c])) v = make_float2(-v.x, -v.y);
      work[pc_spec_index(g, r, kys[c], kxs[c], n3)] = v;
    }
  }
}

// Launches the quadrant block product: k_pc_quadmul (FP32 arithmetic on
// the vector), or with VP_PC_DOT2=1 the dot2 kernel for FP32 applies with
// binary16 blocks.  The dot2 kernel rounds the vector to binary16 too: its
// block product is about 7% faster at the Bragg 10 nm workload, but the
// rounding of the vector makes the preconditioner a nonlinear map, and
// right-preconditioned GMRES there needs 87 inner iterations instead of 67.
template <typename T>
static void launch_quadmul_window(const PcGeom& g, const PcArgs<T>& a, dim3 grid, int threads, int qu,
                                  hipStream_t s) {
  if (qu == 32) hipLaunchKernelGGL((k_pc_quadmul<T, 32>), grid, dim3(threads), 0, s, a, g);
  else if (qu == 8) hipLaunchKernelGGL((k_pc_quadmul<T, 8>), grid, dim3(threads), 0, s, a, g);
  else hipLaunchKernelGGL((k_pc_quadmul<T, 16>), grid, dim3(threads), 0, s, a, g);
  VP_HIP(hipGetLastError());
}

template <typename T>
static void launch_quadmul(const PcGeom& g, const PcArgs<T>& a, dim3 grid, int threads, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    const char* e = std::getenv("VP_PC_DOT2");
    const bool dot2 = e && e[0] == '1';
    if (dot2 && a.hblocks && 3 * g.N <= 255) {
      hipLaunchKernelGGL(k_pc_quadmul_dot2, grid, dim3(threads), 0, s, a, g);
      return;
    }
  }
  // binary16 blocks: columns of block loads in flight per chunk, the
  // preconditioner's tuned window g.qu (VP_PC_QU overrides)
  const char* eq = std::getenv("VP_PC_QU");
  const int qu = eq ? std::atoi(eq) : (g.qu > 0 ? g.qu : 16);
  launch_quadmul_window<T>(g, a, grid, threads, qu, s);
}

int tune_quadmul_window(const PcGeom& g, const PcArgs<float>& a, hipStream_t s) {
  if (!a.hblocks) return 16;
  const dim3 grid(unsigned(int64_t(g.Hx) * g.Hy), 1);
  const int threads = std::min(256, ((3 * g.N + 63) / 64) * 64);
  hipEvent_t e0, e1;
  VP_HIP(hipEventCreate(&e0));
  VP_HIP(hipEventCreate(&e1));
  int best = 16;
  float tbest = 0.f;
  for (int qu : {8, 16, 32}) {
    launch_quadmul_window<float>(g, a, grid, threads, qu, s);   // warm-up
    VP_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 3; ++r) launch_quadmul_window<float>(g, a, grid, threads, qu, s);
    VP_HIP(hipEventRecord(e1, s));
    VP_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    VP_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (qu == 8 || ms < tbest) { tbest = ms; best = qu; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// x pass inverse + gather into the output vector.
template <typename T>
__global__ void __launch_bounds__(512) k_pc_xinv(PcArgs<T> a, PcGeom g) {
  if (a.active && !a.active[blockIdx.y]) return;
  C<T>* __restrict__ Y = a.y + int64_t(blockIdx.y) * a.ystride;
  const C<T>* __restrict__ work = a.work + int64_t(blockIdx.y) * a.wstride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);
  const int L = g.L, M = g.M, N = g.N, B = g.B, ld = g.ld;
  C<T>* W1 = W0 + B * ld;
  const int nmb = (M + B - 1) / B;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int mb = blk % nmb, rest = blk / nmb;
  const int n = rest % N, c = rest / N;
  const int m0 = mb * B;
  const int64_t MN = int64_t(M) * N;
  for (int e = threadIdx.x; e < B * L; e += blockDim.x) {
    const int kx = e / B, bb = e - kx * B, m = m0 + bb;
    W0[bb * ld + kx] = (m < M) ? work[(int64_t(kx) * 3 + c) * MN + int64_t(n) * M + m] : czero<T>();
  }
  __syncthreads();
  const C<T>* R = lds_fft<T, true>(W0, W1, g.px, B, LdsLayout{1, ld, 0, 1}, a.twx);
  for (int e = threadIdx.x; e < B * L; e += blockDim.x) {
    const int bb = e / L, l = e - bb * L, m = m0 + bb;
    if (m < M) {
      const int64_t vox = l + int64_t(L) * (m + int64_t(M) * n);
      int64_t j;
      if (g.structured) {
        const int p = a.pos[vox];
        j = p < 0 ? -1 : p + int64_t(c) * g.nS;
      } else {
        j = a.pos[int64_t(c) * g.V + vox];
      }
      if (j >= 0) Y[pc_outer(a, j)] = R[bb * ld + l];
    }
  }
}

// 1-level: block product over the whole 3MN x-frequency slice.
template <typename T>
__global__ void __launch_bounds__(256) k_pc1_block(PcArgs<T> a, PcGeom g) {
  if (a.active && !a.active[blockIdx.y]) return;
  C<T>* __restrict__ work = a.work + int64_t(blockIdx.y) * a.wstride;
  C<T>* __restrict__ out = a.work2 + int64_t(blockIdx.y) * a.wstride;
  const C<T>* __restrict__ blocks = a.blocks[blockIdx.y];
  const int64_t n3 = 3LL * g.M * g.N;
  const int kx = blockIdx.x;
  const C<T>* Bk = blocks + int64_t(kx) * n3 * n3;
  const C<T>* v = work + int64_t(kx) * n3;
  for (int64_t r = threadIdx.x + int64_t(blockIdx.z) * blockDim.x; r < n3; r += int64_t(gridDim.z) * blockDim.x) {
    C<T> acc = czero<T>();
    for (int64_t q = 0; q < n3; ++q) acc = cadd(acc, cmul(Bk[q * n3 + r], v[q]));
    out[int64_t(kx) * n3 + r] = acc;
  }
}

template <typename T>
__global__ void k_copy_vec(const C<T>* __restrict__ x, C<T>* __restrict__ y, int64_t n, int64_t xstride,
                           int64_t ystride, const int* __restrict__ active) {
  if (active && !active[blockIdx.y]) return;
  const C<T>* X = x + int64_t(blockIdx.y) * xstride;
  C<T>* Y = y + int64_t(blockIdx.y) * ystride;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    Y[i] = X[i];
}

// ---------------------------------------------------------------------------
// Quadrant form over pencil-major spectra X[c][n][m][kx] (g.pct): the x
// passes take one pencil (c, n, m) per workgroup with a compile-time
// length-L FFT contiguous in memory; the row passes transform NB
// consecutive kx of one (c, n) along m in place, read and written
// b-fastest; k_pc_quadmul addresses the same layout.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ int64_t pc_unknown(const PcArgs<T>& a, const PcGeom& g, int64_t vox, int c) {
  if (g.structured) {
    const int p = a.pos[vox];
    return p < 0 ? -1 : p + int64_t(c) * g.nS;
  }
  return a.pos[int64_t(c) * g.V + vox];
}

template <typename T, int NT, int... Rs>
__global__ void __launch_bounds__(NT) k_pc_xfwd_ct(PcArgs<T> a, PcGeom g, int twlds) {
  constexpr int L = (Rs * ...);
  if (a.active && !a.active[blockIdx.y]) return;
  const C<T>* __restrict__ X = a.x + int64_t(blockIdx.y) * a.xstride;
  C<T>* __restrict__ work = a.work + int64_t(blockIdx.y) * a.wstride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);     // one buffer: the stages run in place
  const int M = g.M, N = g.N;
  const int p = xcd_remap(blockIdx.x, gridDim.x);   // pencil (c, n, m), m fastest
  const int cn = p / M, m = p - cn * M;
  const int c = cn / N, n = cn - c * N;
  auto ld = [&](int, int l) {
    const int64_t j = pc_unknown(a, g, l + int64_t(L) * (m + int64_t(M) * n), c);
    return j >= 0 ? X[pc_outer(a, j)] : czero<T>();
  };
  const BufRes Wo = buf_res(work + int64_t(p) * L, uint32_t(L) * uint32_t(sizeof(C<T>)));
  auto st = [&](int, int k, C<T> v) { buf_store(Wo, uint32_t(k), v); };
  if (twlds) {   // stage twiddles from two short FP64 factor tables in LDS
    const TwSplit<T> tws = TwSplit<T>::fill(W0 + L, L);
    __syncthreads();
    ct_fft_ip<T, false, L, 1, 1, NT, 1, true, false, false, false, Rs...>(W0, tws, 1, ld, st);
  } else {
    ct_fft_ip<T, false, L, 1, 1, NT, 1, true, false, false, false, Rs...>(W0, a.twx, 1, ld, st);
  }
}

template <typename T, int NT, int... Rs>
__global__ void __launch_bounds__(NT) k_pc_xinv_ct(PcArgs<T> a, PcGeom g, int twlds) {
  constexpr int L = (Rs * ...);
  if (a.active && !a.active[blockIdx.y]) return;
  C<T>* __restrict__ Y = a.y + int64_t(blockIdx.y) * a.ystride;
  const C<T>* __restrict__ work = a.work + int64_t(blockIdx.y) * a.wstride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);     // one buffer: the stages run in place
  const int M = g.M, N = g.N;
  const int p = xcd_remap(blockIdx.x, gridDim.x);
  const int cn = p / M, m = p - cn * M;
  const int c = cn / N, n = cn - c * N;
  const BufRes Wi = buf_res(work + int64_t(p) * L, uint32_t(L) * uint32_t(sizeof(C<T>)));
  auto ld = [&](int, int k) { return buf_load<T>(Wi, uint32_t(k)); };
  auto none = [](int, int, C<T>) {};
  const C<T>* R;
  if (twlds) {
    const TwSplit<T> tws = TwSplit<T>::fill(W0 + L, L);
    __syncthreads();
    R = ct_fft_ip<T, true, L, 1, 1, NT, 1, true, false, false, true, Rs...>(W0, tws, 1, ld, none);
  } else {
    R = ct_fft_ip<T, true, L, 1, 1, NT, 1, true, false, false, true, Rs...>(W0, a.twx, 1, ld, none);
  }
  for (int l = threadIdx.x; l < L; l += NT) {
    const int64_t j = pc_unknown(a, g, l + int64_t(L) * (m + int64_t(M) * n), c);
    if (j >= 0) Y[pc_outer(a, j)] = R[l];
  }
}

template <typename T, bool INV, int NB, int... Rs>
__global__ void __launch_bounds__(256) k_pc_rows_ct(PcArgs<T> a, PcGeom g) {
  constexpr int MM = (Rs * ...);
  constexpr int LD = NB + 1;
  if (a.active && !a.active[blockIdx.y]) return;
  C<T>* __restrict__ work = a.work + int64_t(blockIdx.y) * a.wstride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);
  C<T>* W1 = W0 + MM * LD;
  C<T>* tw = W1 + MM * LD;
  for (int e = threadIdx.x; e < MM; e += 256) tw[e] = a.twy[e];
  const int L = g.L;
  const int nkb = (L + NB - 1) / NB;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int cn = blk / nkb, kb = blk - cn * nkb;
  const int kx0 = kb * NB;
  const int nb = min(NB, L - kx0);
  // (m, kx0 + b) at m*L + b from the row block's origin
  const BufRes X = buf_res(work + int64_t(cn) * MM * L + kx0, uint32_t((MM - 1) * L + nb) * uint32_t(sizeof(C<T>)));
  __syncthreads();
  auto ld = [&](int b, int m) { return b < nb ? buf_load<T>(X, uint32_t(m * L + b)) : czero<T>(); };
  auto st = [&](int b, int m, C<T> v) {
    if (b < nb) buf_store(X, uint32_t(m * L + b), v);
  };
  ct_fft<T, INV, MM, NB, LD, 256, 1, true, true, true, false, Rs...>(W0, W1, tw, 1, ld, st);
}

// Pencil-major spectra X[q][m][kx] (work) <-> q-contiguous Y[kx][m][q]
// (work2), q = c*N + n < 3N: 32 x 32 tiles of one m through LDS, read and
// written as 256-byte runs.  Grid: (kx tiles * q tiles, M, problems).
template <typename T, bool TO_Q>
__global__ void __launch_bounds__(256) k_pc_qlay(PcArgs<T> a, PcGeom g) {
  const int p = blockIdx.z;
  if (a.active && !a.active[p]) return;
  const int L = g.L, M = g.M, n3 = 3 * g.N;
  const int xt = (L + 31) / 32;
  const int kx0 = int(blockIdx.x % unsigned(xt)) * 32, q0 = int(blockIdx.x / unsigned(xt)) * 32;
  const int m = blockIdx.y;
  C<T>* __restrict__ X = a.work + int64_t(p) * a.wstride;
  C<T>* __restrict__ Y = a.work2 + int64_t(p) * a.wstride;
  __shared__ C<T> tile[32][33];
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;   // 32 x 8 threads
  if (TO_Q) {
    for (int r = ly; r < 32; r += 8) {
      const int q = q0 + r, kx = kx0 + lx;
      if (q < n3 && kx < L) tile[r][lx] = X[(int64_t(q) * M + m) * L + kx];
    }
    __syncthreads();
    for (int r = ly; r < 32; r += 8) {
      const int kx = kx0 + r, q = q0 + lx;
      if (q < n3 && kx < L) Y[(int64_t(kx) * M + m) * n3 + q] = tile[lx][r];
    }
  } else {
    for (int r = ly; r < 32; r += 8) {
      const int kx = kx0 + r, q = q0 + lx;
      if (q < n3 && kx < L) tile[lx][r] = Y[(int64_t(kx) * M + m) * n3 + q];
    }
    __syncthreads();
    for (int r = ly; r < 32; r += 8) {
      const int q = q0 + r, kx = kx0 + lx;
      if (q < n3 && kx < L) X[(int64_t(q) * M + m) * L + kx] = tile[r][lx];
    }
  }
}

// One line of length R1*R2 (R2 = 1: a single radix) in registers:
// x[n], n = R2*n1 + n2, becomes X[k1 + R1*k2] (R1-point DFTs over n1, twiddles
// w^(n2*k1), R2-point DFTs over n2).
template <typename T, bool INV, int R1, int R2>
__device__ __forceinline__ void dft_line(C<T>* v) {
  if constexpr (R2 == 1) {
    dft<T, R1, INV>(v);
  } else {
    constexpr int MM = R1 * R2;
    constexpr RootTable<MM> rt{};
    C<T> y[MM];   // y[n2 * R1 + k1]
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      C<T> u[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) u[n1] = v[R2 * n1 + n2];
      dft<T, R1, INV>(u);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const int e = (n2 * k1) % MM;
        C<T> w;
        w.x = T(rt.c[e]);
        w.y = INV ? T(rt.s[e]) : T(-rt.s[e]);
        y[n2 * R1 + k1] = e == 0 ? u[k1] : cmul(u[k1], w);
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      C<T> b[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[n2 * R1 + k1];
      dft<T, R2, INV>(b);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
    }
  }
}

// Row transforms (length M = R1*R2) fused with the change of layout around
// the block product: forward X[q][m][kx] (work) -> Y[kx][m][q] (work2),
// inverse Y -> X.  A workgroup takes QB rows q by KB columns kx, one line
// per thread.  The pencil-major side is read or written straight from the
// registers (a wavefront touches runs of KB consecutive kx); the q-contiguous
// side goes through an LDS tile of the workgroup's lines, read or written as
// runs of QB consecutive q.
template <typename T, bool INV, int R1, int R2>
__global__ void __launch_bounds__(256) k_pc_rows_q(PcArgs<T> a, PcGeom g) {
  constexpr int MM = R1 * R2, QB = 16, KB = 16, LDK = KB + 1;
  if (a.active && !a.active[blockIdx.y]) return;
  C<T>* __restrict__ X = a.work + int64_t(blockIdx.y) * a.wstride;
  C<T>* __restrict__ Y = a.work2 + int64_t(blockIdx.y) * a.wstride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* S = reinterpret_cast<C<T>*>(smem_raw);   // S[(m * QB + ql) * LDK + kl]
  const int L = g.L, n3 = 3 * g.N;
  const int nkt = (L + KB - 1) / KB;
  const int kx0 = int(blockIdx.x % unsigned(nkt)) * KB, q0 = int(blockIdx.x / unsigned(nkt)) * QB;
  const int t = threadIdx.x;
  const int ql = t / KB, kl = t % KB;
  const int q = q0 + ql, kx = kx0 + kl;
  const bool mine = q < n3 && kx < L;
  C<T> v[MM];
  if (!INV) {
#pragma unroll
    for (int m = 0; m < MM; ++m) v[m] = mine ? X[(int64_t(q) * MM + m) * L + kx] : czero<T>();
    dft_line<T, false, R1, R2>(v);
#pragma unroll
    for (int m = 0; m < MM; ++m) S[(m * QB + ql) * LDK + kl] = v[m];
    __syncthreads();
    for (int e = t; e < MM * QB * KB; e += 256) {   // q fastest
      const int eq = e % QB, r = e / QB, m = r % MM, ek = r / MM;
      if (q0 + eq < n3 && kx0 + ek < L)
        Y[(int64_t(kx0 + ek) * MM + m) * n3 + q0 + eq] = S[(m * QB + eq) * LDK + ek];
    }
  } else {
    // lines taken q-fastest, so a wavefront reads runs of QB consecutive q
    const int iq = t % QB, ik = t / QB;
    const bool mine_q = q0 + iq < n3 && kx0 + ik < L;
#pragma unroll
    for (int m = 0; m < MM; ++m)
      v[m] = mine_q ? Y[(int64_t(kx0 + ik) * MM + m) * n3 + q0 + iq] : czero<T>();
    dft_line<T, true, R1, R2>(v);
#pragma unroll
    for (int m = 0; m < MM; ++m) S[(m * QB + iq) * LDK + ik] = v[m];
    __syncthreads();
    for (int e = t; e < MM * QB * KB; e += 256) {   // kx fastest
      const int ek = e % KB, r = e / KB, eq = r % QB, m = r / QB;
      if (q0 + eq < n3 && kx0 + ek < L)
        X[(int64_t(q0 + eq) * MM + m) * L + kx0 + ek] = S[(m * QB + eq) * LDK + ek];
    }
  }
}

// max over entries of max(|re|, |im|) (non-negative floats order as their bits)
__global__ void k_absmax_c64(const float2* __restrict__ a, int64_t n, unsigned* __restrict__ out) {
  float m = 0.f;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    m = fmaxf(m, fmaxf(fabsf(a[i].x), fabsf(a[i].y)));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

__global__ void k_to_half_pairs(const float2* __restrict__ a, int64_t n, float inv, uint32_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const _Float16 re = _Float16(a[i].x * inv), im = _Float16(a[i].y * inv);
    out[i] = uint32_t(__builtin_bit_cast(uint16_t, re)) | (uint32_t(__builtin_bit_cast(uint16_t, im)) << 16);
  }
}

void make_half_blocks(PrecPrec<float>& P, hipStream_t s) {
  const int64_t n = int64_t(P.blocks.bytes / sizeof(float2));
  DevBuf mx(sizeof(unsigned));
  VP_HIP(hipMemsetAsync(mx.p, 0, sizeof(unsigned), s));
  const unsigned nb = unsigned(std::min<int64_t>((n + 255) / 256, 16384));
  hipLaunchKernelGGL(k_absmax_c64, dim3(nb), dim3(256), 0, s, P.blocks.as<float2>(), n, mx.as<unsigned>());
  unsigned bits = 0;
  VP_HIP(hipMemcpyAsync(&bits, mx.p, sizeof(bits), hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  float m;
  std::memcpy(&m, &bits, sizeof(m));
  // entries land in [-1, 1] x 2^15: well inside binary16's normal range
  P.hscale = m > 0.f ? m / 32768.f : 1.f;
  P.hblocks.alloc(size_t(n) * sizeof(uint32_t));
  hipLaunchKernelGGL(k_to_half_pairs, dim3(nb), dim3(256), 0, s, P.blocks.as<float2>(), n, 1.f / P.hscale,
                     P.hblocks.as<uint32_t>());
  VP_HIP(hipGetLastError());
  const uint32_t* hp = P.hblocks.as<uint32_t>();
  P.hblockptr.alloc(sizeof(hp));
  VP_HIP(hipMemcpyAsync(P.hblockptr.p, &hp, sizeof(hp), hipMemcpyHostToDevice, s));
  P.hscaleptr.alloc(sizeof(float));
  VP_HIP(hipMemcpyAsync(P.hscaleptr.p, &P.hscale, sizeof(float), hipMemcpyHostToDevice, s));
  VP_HIP(hipStreamSynchronize(s));
}

// Quadrant form over pencil-major spectra (g.pct).
template <typename T>
static void launch_prec_pct(const PcGeom& g, const PcArgs<T>& a, int nprob, hipStream_t s) {
  constexpr bool F64 = sizeof(T) == 8;
  const unsigned npen = unsigned(3 * g.N * g.M);
  const size_t ldsx = size_t(g.L) * sizeof(C<T>) + TwSplit<T>::entries(g.L) * sizeof(double2);
  const int twlds = [] {   // VP_XTW_LDS=0: stage twiddles from the global table (read per launch)
    const char* e = std::getenv("VP_XTW_LDS");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  auto xpass = [&](bool inv) {
    bool done = false;
#define VP_PCX(SL, NTF, FR, NTI, IR)                                                                      \
  if (!done && g.L == SL) {                                                                             \
    if (inv) {                                                                                          \
      const void* fn = reinterpret_cast<const void*>(&k_pc_xinv_ct<T, NTI, VP_UNPAREN IR>);             \
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsx));             \
      hipLaunchKernelGGL((k_pc_xinv_ct<T, NTI, VP_UNPAREN IR>), dim3(npen, nprob), dim3(NTI), ldsx, s, a, g, twlds); \
    } else {                                                                                            \
      const void* fn = reinterpret_cast<const void*>(&k_pc_xfwd_ct<T, NTF, VP_UNPAREN FR>);             \
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsx));             \
      hipLaunchKernelGGL((k_pc_xfwd_ct<T, NTF, VP_UNPAREN FR>), dim3(npen, nprob), dim3(NTF), ldsx, s, a, g, twlds); \
    }                                                                                                   \
    done = true;                                                                                        \
  }
    if constexpr (F64) {
      VP_XCT64_LIST(VP_PCX)
    } else {
      VP_XCT32_LIST(VP_PCX)
    }
#undef VP_PCX
    if (!done) throw VpError{ERR_UNSUPPORTED, "preconditioner: no compile-time x pass for this length"};
  };
  constexpr int NB = 16;
  const unsigned nrowblk = unsigned(3 * g.N * ((g.L + NB - 1) / NB));
  auto rows = [&](bool inv) {
    bool done = false;
#define VP_PCR(NL, ...)                                                                                 \
  if (!done && g.M == NL) {                                                                             \
    const size_t lds = (2 * size_t(NL) * (NB + 1) + NL) * sizeof(C<T>);                                 \
    if (inv)                                                                                            \
      hipLaunchKernelGGL((k_pc_rows_ct<T, true, NB, __VA_ARGS__>), dim3(nrowblk, nprob), dim3(256), lds, s, a, g); \
    else                                                                                                \
      hipLaunchKernelGGL((k_pc_rows_ct<T, false, NB, __VA_ARGS__>), dim3(nrowblk, nprob), dim3(256), lds, s, a, g); \
    done = true;                                                                                        \
  }
    VP_PCROW_LIST(VP_PCR)
#undef VP_PCR
    if (!done) throw VpError{ERR_UNSUPPORTED, "preconditioner: no compile-time row pass for this length"};
  };
  const dim3 gq(unsigned(int64_t(g.Hx) * g.Hy), nprob);
  const int tq = std::min(256, ((3 * g.N + 63) / 64) * 64);
  {
    ProfScope ps(s, KID_PC_XFWD);
    xpass(false);
  }
  // q-contiguous spectra for the block product: the row passes write and read
  // them directly where the fused kernel fits (length with at most two
  // radices, LDS within 64 KiB), else separate LDS transposes around the
  // pencil-major row passes
  const dim3 gt(unsigned(((g.L + 31) / 32) * ((3 * g.N + 31) / 32)), unsigned(g.M), unsigned(nprob));
  const dim3 gr(unsigned(((g.L + 15) / 16) * ((3 * g.N + 15) / 16)), unsigned(nprob));
  const size_t ldsq = size_t(g.M) * 16 * 17 * sizeof(C<T>);
  auto rows_q = [&](bool inv) {
    if (!g.qlay || ldsq > 65536) return false;
    bool done = false;
#define VP_PCQ(NL, R1, R2)                                                                              \
  if (!done && g.M == NL) {                                                                             \
    static_assert(NL == R1 * R2, "row length is the product of its radices");                           \
    const void* fn = inv ? reinterpret_cast<const void*>(&k_pc_rows_q<T, true, R1, R2>)                 \
                         : reinterpret_cast<const void*>(&k_pc_rows_q<T, false, R1, R2>);               \
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsq));               \
    if (inv)                                                                                            \
      hipLaunchKernelGGL((k_pc_rows_q<T, true, R1, R2>), gr, dim3(256), ldsq, s, a, g);                 \
    else                                                                                                \
      hipLaunchKernelGGL((k_pc_rows_q<T, false, R1, R2>), gr, dim3(256), ldsq, s, a, g);                \
    done = true;                                                                                        \
  }
    VP_PCQ(11, 11, 1) VP_PCQ(22, 2, 11) VP_PCQ(10, 2, 5) VP_PCQ(12, 3, 4) VP_PCQ(20, 4, 5) VP_PCQ(24, 3, 8)
    VP_PCQ(44, 4, 11) VP_PCQ(16, 4, 4) VP_PCQ(32, 4, 8)
#undef VP_PCQ
    return done;
  };
  {
    ProfScope ps(s, KID_PC_YROWS);
    if (!rows_q(false)) {
      rows(false);
      if (g.qlay) hipLaunchKernelGGL((k_pc_qlay<T, true>), gt, dim3(256), 0, s, a, g);
    }
  }
  {
    ProfScope ps(s, KID_PC_BLOCK);
    launch_quadmul<T>(g, a, gq, tq, s);
  }
  {
    ProfScope ps(s, KID_PC_YROWS);
    if (!rows_q(true)) {
      if (g.qlay) hipLaunchKernelGGL((k_pc_qlay<T, false>), gt, dim3(256), 0, s, a, g);
      rows(true);
    }
  }
  {
    ProfScope ps(s, KID_PC_XINV);
    xpass(true);
  }
  VP_HIP(hipGetLastError());
}

template <typename T>
void launch_prec(const PcGeom& g, const PcArgs<T>& a, int nprob, hipStream_t s) {
  if (a.sub && g.n_outer > 0 && !a.skip_copy) {
    const int blocks = int(std::min<int64_t>((g.n_outer + 255) / 256, 4096));
    hipLaunchKernelGGL(k_copy_vec<T>, dim3(blocks, nprob), dim3(256), 0, s, a.x, a.y, g.n_outer, a.xstride,
                       a.ystride, a.active);
    VP_HIP(hipGetLastError());
  }
  const int nmb = (g.M + g.B - 1) / g.B;
  const dim3 gx(nmb * g.N * 3, nprob);
  // the geometry sizes LDS for complex128; complex64 launches need half
  const size_t lds1 = g.lds1 * sizeof(C<T>) / 16, lds2 = g.lds2 * sizeof(C<T>) / 16;
  const void* f1 = reinterpret_cast<const void*>(&k_pc_xfwd<T>);
  const void* f3 = reinterpret_cast<const void*>(&k_pc_xinv<T>);
  (void)hipFuncSetAttribute(f1, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds1));
  (void)hipFuncSetAttribute(f3, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds1));
  if (g.level == 2 && g.pquad && g.pct) {
    launch_prec_pct<T>(g, a, nprob, s);
    return;
  }
  {
    ProfScope ps(s, KID_PC_XFWD);
    hipLaunchKernelGGL(k_pc_xfwd<T>, gx, dim3(g.tx), lds1, s, a, g);
  }
  VP_HIP(hipGetLastError());
  if (g.level == 2 && g.pquad) {
    const size_t ldr = g.lds_row * sizeof(C<T>) / 16;
    const void* fr0 = reinterpret_cast<const void*>(&k_pc_yrows<T, false>);
    const void* fr1 = reinterpret_cast<const void*>(&k_pc_yrows<T, true>);
    (void)hipFuncSetAttribute(fr0, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldr));
    (void)hipFuncSetAttribute(fr1, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldr));
    const int64_t nrows = int64_t(g.L) * 3 * g.N;
    const dim3 gr(unsigned((nrows + g.Brow - 1) / g.Brow), nprob);
    const dim3 gq(unsigned(int64_t(g.Hx) * g.Hy), nprob);
    const int tq = std::min(256, ((3 * g.N + 63) / 64) * 64);
    {
      // row transforms and the block products timed separately
      {
        ProfScope ps(s, KID_PC_YROWS);
        hipLaunchKernelGGL((k_pc_yrows<T, false>), gr, dim3(256), ldr, s, a, g);
      }
      {
        ProfScope ps(s, KID_PC_BLOCK);
        launch_quadmul<T>(g, a, gq, tq, s);
      }
      ProfScope ps(s, KID_PC_YROWS);
      hipLaunchKernelGGL((k_pc_yrows<T, true>), gr, dim3(256), ldr, s, a, g);
    }
    VP_HIP(hipGetLastError());
    ProfScope ps(s, KID_PC_XINV);
    hipLaunchKernelGGL(k_pc_xinv<T>, gx, dim3(g.tx), lds1, s, a, g);
  } else if (g.level == 2 && g.psplit) {
    const size_t ldr = g.lds_row * sizeof(C<T>) / 16, ldt = g.lds_tile * sizeof(C<T>) / 16;
    const void* fr0 = reinterpret_cast<const void*>(&k_pc_yrows<T, false>);
    const void* fr1 = reinterpret_cast<const void*>(&k_pc_yrows<T, true>);
    const void* fb = reinterpret_cast<const void*>(&k_pc_blockmul<T>);
    (void)hipFuncSetAttribute(fr0, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldr));
    (void)hipFuncSetAttribute(fr1, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldr));
    (void)hipFuncSetAttribute(fb, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldt));
    const int64_t nrows = int64_t(g.L) * 3 * g.N;
    const dim3 gr(unsigned((nrows + g.Brow - 1) / g.Brow), nprob);
    const dim3 gb(unsigned(int64_t(g.L) * ((g.M + g.TK - 1) / g.TK)), nprob);
    {
      {
        ProfScope ps(s, KID_PC_YROWS);
        hipLaunchKernelGGL((k_pc_yrows<T, false>), gr, dim3(256), ldr, s, a, g);
      }
      {
        ProfScope ps(s, KID_PC_BLOCK);
        hipLaunchKernelGGL(k_pc_blockmul<T>, gb, dim3(256), ldt, s, a, g);
      }
      ProfScope ps(s, KID_PC_YROWS);
      hipLaunchKernelGGL((k_pc_yrows<T, true>), gr, dim3(256), ldr, s, a, g);
    }
    VP_HIP(hipGetLastError());
    ProfScope ps(s, KID_PC_XINV);
    hipLaunchKernelGGL(k_pc_xinv<T>, gx, dim3(g.tx), lds1, s, a, g);
  } else if (g.level == 2) {
    const void* f2 = reinterpret_cast<const void*>(&k_pc_plane<T>);
    (void)hipFuncSetAttribute(f2, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds2));
    {
      ProfScope ps(s, KID_PC_PLANE);
      hipLaunchKernelGGL(k_pc_plane<T>, dim3(g.L, nprob), dim3(g.t2), lds2, s, a, g);
    }
    VP_HIP(hipGetLastError());
    ProfScope ps(s, KID_PC_XINV);
    hipLaunchKernelGGL(k_pc_xinv<T>, gx, dim3(g.tx), lds1, s, a, g);
  } else {
    const int64_t n3 = 3LL * g.M * g.N;
    const int zs = int(std::min<int64_t>((n3 + 255) / 256, 64));
    hipLaunchKernelGGL(k_pc1_block<T>, dim3(g.L, nprob, zs), dim3(256), 0, s, a, g);
    VP_HIP(hipGetLastError());
    PcArgs<T> b = a;
    b.work = a.work2;
    hipLaunchKernelGGL(k_pc_xinv<T>, gx, dim3(g.tx), lds1, s, b, g);
  }
  VP_HIP(hipGetLastError());
}
template void launch_prec<float>(const PcGeom&, const PcArgs<float>&, int, hipStream_t);
template void launch_prec<double>(const PcGeom&, const PcArgs<double>&, int, hipStream_t);

// ----------------------------------------------------------------------------
// Setup kernels (FP64)
// ----------------------------------------------------------------------------
__constant__ int c_pc_even_x[6] = {1, 0, 0, 1, 1, 1};
__constant__ int c_pc_comp[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

// ck[k][kx][m][n] = sum_i chan_x(A_k)(i) w_L^{i kx}, A_k = -Mc*G_k (+ Gram*Mr0 at origin
// for xx, yy, zz); chan_x per level_1_parallel_func.m:12-21.
__global__ void k_pc_chan_x(const double2* __restrict__ G, double2* __restrict__ ck, int L, int M, int N,
                            double2 mc, double diag, const double2* __restrict__ twL) {
  const int64_t total = int64_t(6) * L * M * N;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = e;
    const int kx = int(r % L); r /= L;
    const int m = int(r % M); r /= M;
    const int n = int(r % N); r /= N;
    const int k = int(r);
    const int64_t V = int64_t(L) * M * N;
    const int64_t base = int64_t(L) * (m + int64_t(M) * n) + V * k;
    auto A = [&](int i) -> double2 {
      double2 g = G[base + i];
      double2 a = make_double2(-(mc.x * g.x - mc.y * g.y), -(mc.x * g.y + mc.y * g.x));
      if (i == 0 && m == 0 && n == 0 && (k == 0 || k == 3 || k == 5)) a.x += diag;
      return a;
    };
    const double sg = c_pc_even_x[k] ? 1.0 : -1.0;
    double2 acc = A(0);
    int64_t ph = kx;
    for (int i = 1; i < L; ++i, ph += kx) {
      if (ph >= L) ph -= L;
      const double2 ai = A(i), aj = A(L - i);
      const double wa = double(L - i) / L, wb = sg * double(i) / L;
      const double2 cv = make_double2(wa * ai.x + wb * aj.x, wa * ai.y + wb * aj.y);
      const double2 w = twL[ph];
      acc.x += cv.x * w.x - cv.y * w.y;
      acc.y += cv.x * w.y + cv.y * w.x;
    }
    ck[e] = acc;   // layout [k][n][m][kx]
  }
}

// Same coefficients, one workgroup per line (k, n, m): the Chan-weighted line
// c_i (i < L) is formed once in LDS and its length-L DFT runs in two levels,
// L = L1*L2, i = i1 + L1*i2, kx = k2 + L2*k1:
//   T[k2*L1 + i1] = w^{i1 k2} sum_{i2} c_{i1 + L1 i2} w^{L1 i2 k2},
//   ck[kx]        = sum_{i1} T[k2*L1 + i1] w^{L2 i1 k1},
// L*(L1 + L2) products instead of L^2 (w = twL[1], indices reduced mod L).
__global__ void __launch_bounds__(256) k_pc_chan_x2(const double2* __restrict__ G, double2* __restrict__ ck, int L,
                                                    int M, int N, int L1, double2 mc, double diag,
                                                    const double2* __restrict__ twL) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* c = reinterpret_cast<double2*>(smem_raw);
  double2* T = c + L;
  const int L2 = L / L1;
  int r = blockIdx.x;
  const int m = r % M; r /= M;
  const int n = r % N;
  const int k = r / N;
  const int64_t V = int64_t(L) * M * N;
  const int64_t base = int64_t(L) * (m + int64_t(M) * n) + V * k;
  const double sg = c_pc_even_x[k] ? 1.0 : -1.0;
  const bool dg = m == 0 && n == 0 && (k == 0 || k == 3 || k == 5);
  auto A = [&](int i) -> double2 {
    const double2 g = G[base + i];
    double2 a = make_double2(-(mc.x * g.x - mc.y * g.y), -(mc.x * g.y + mc.y * g.x));
    if (i == 0 && dg) a.x += diag;
    return a;
  };
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    double2 v = A(i);
    if (i > 0) {
      const double2 aj = A(L - i);
      const double wa = double(L - i) / L, wb = sg * double(i) / L;
      v = make_double2(wa * v.x + wb * aj.x, wa * v.y + wb * aj.y);
    }
    c[i] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < L; e += blockDim.x) {
    const int k2 = e / L1, i1 = e - k2 * L1;
    const int step = L1 * k2;   // < L
    double2 acc = make_double2(0.0, 0.0);
    int ph = 0;
    for (int i2 = 0; i2 < L2; ++i2) {
      const double2 v = c[i1 + L1 * i2], w = twL[ph];
      acc.x += v.x * w.x - v.y * w.y;
      acc.y += v.x * w.y + v.y * w.x;
      ph += step;
      if (ph >= L) ph -= L;
    }
    const double2 w = twL[(i1 * k2) % L];
    T[e] = make_double2(acc.x * w.x - acc.y * w.y, acc.x * w.y + acc.y * w.x);
  }
  __syncthreads();
  double2* out = ck + (int64_t(k) * N * M + int64_t(n) * M + m) * L;   // layout [k][n][m][kx]
  for (int kx = threadIdx.x; kx < L; kx += blockDim.x) {
    const int k1 = kx / L2, k2 = kx - k1 * L2;
    const int step = L2 * k1;   // < L
    double2 acc = make_double2(0.0, 0.0);
    int ph = 0;
    for (int i1 = 0; i1 < L1; ++i1) {
      const double2 v = T[k2 * L1 + i1], w = twL[ph];
      acc.x += v.x * w.x - v.y * w.y;
      acc.y += v.x * w.y + v.y * w.x;
      ph += step;
      if (ph >= L) ph -= L;
    }
    out[kx] = acc;
  }
}

// Block (kx, ky) entry [(c, nt), (d, ns)] = sum_t chan_y(t) w_M^{t ky}.  The
// M x M y-Toeplitz block of component k at z-offset (nt, ns) has first
// column col(t) and first row row(t) built from v(t) = ck[k][|nt-ns|][t][kx]
// with the sign conventions of level_2_parallel_func.m:94-162 (toeplitz
// arguments for xy and yz, the triu completions for xz and yz).
__device__ __forceinline__ double2 neg2(double2 a) { return make_double2(-a.x, -a.y); }

__device__ __forceinline__ double2 pc_col(int k, int nt, int ns, int t, double2 v) {
  if (k == 0 || k == 3 || k == 5) return v;
  if (k == 1) return t == 0 ? neg2(v) : v;
  if (k == 2) {
    if (nt < ns) return neg2(v);
    if (nt > ns) return v;
    return t == 0 ? neg2(v) : v;
  }
  if (nt < ns) return t == 0 ? v : neg2(v);   // k == 4
  return v;
}

__device__ __forceinline__ double2 pc_row(int k, int nt, int ns, int t, double2 v) {
  if (k == 0 || k == 3 || k == 5) return v;
  if (k == 1) return neg2(v);
  if (k == 2) return (nt > ns) ? v : neg2(v);
  if (nt > ns) return t == 0 ? v : neg2(v);   // k == 4
  return v;
}

// The y-Chan transform of the 2-level blocks by FFT, for long y axes.  On
// one line (k, dz, kx) of the x-Chan coefficients v(t) = ck[k][dz][t][kx],
// the Chan-weighted sum of k_pc_blocks is, for t >= 1, a column term with
// weight (M-t)/M on v(t) and a row term with weight t/M on v(M-t); both are
// the transform of a(t) = (M-t)/M v(t) (a(0) = 0):
//   A(ky) = sum_{t>=1} a(t) w^{t ky},  sum_{t>=1} t/M v(M-t) w^{t ky} = A(-ky),
// so an entry is s0 v(0) + s_col A(ky) + s_row A(-ky) with the component's
// signs away from t = 0 (pc_col, pc_row).  One workgroup per line; the line
// is read and written with stride L (setup only).
__global__ void __launch_bounds__(256) k_pc_chan_y_fft(const double2* __restrict__ ck, double2* __restrict__ out,
                                                       int L, int M, int kxn, FftPlan plan,
                                                       const double2* __restrict__ twM) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* W0 = reinterpret_cast<double2*>(smem_raw);
  double2* W1 = W0 + M;
  const int64_t line = blockIdx.x;
  const int kx = int(line % kxn);
  const int64_t kd = line / kxn;   // k * N + dz
  const double2* v = ck + kd * int64_t(M) * L + kx;
  double2* o = out + kd * int64_t(M) * L + kx;
  for (int t = threadIdx.x; t < M; t += blockDim.x) {
    const double w = t == 0 ? 0.0 : double(M - t) / M;
    const double2 a = v[int64_t(t) * L];
    W0[t] = make_double2(w * a.x, w * a.y);
  }
  __syncthreads();
  const LdsLayout lay{1, M, 0, 1};
  const double2* R = lds_fft<double, false>(W0, W1, plan, 1, lay, twM);
  for (int t = threadIdx.x; t < M; t += blockDim.x) o[int64_t(t) * L] = R[t];
}

// Block entries from the y-transformed lines (see k_pc_chan_y_fft); same
// output and face handling as k_pc_blocks.
__global__ void k_pc_blocks_fft(const double2* __restrict__ ck, const double2* __restrict__ ay,
                                double2* __restrict__ blk, int L, int M, int N, int kxn, int kyn,
                                const double2* __restrict__ face, double gram) {
  const int n3 = 3 * N;
  const int64_t total = int64_t(kxn) * kyn * n3 * n3;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = e;
    const int row = int(r % n3); r /= n3;
    const int colq = int(r % n3); r /= n3;
    const int ky = int(r % kyn); r /= kyn;
    const int kx = int(r);
    const int c = row / N, nt = row - c * N;
    const int d = colq / N, ns = colq - d * N;
    const int k = c_pc_comp[c][d];
    const int dz = nt > ns ? nt - ns : ns - nt;
    const int64_t base = (int64_t(k) * N + dz) * int64_t(M) * L + kx;
    double2 acc = pc_col(k, nt, ns, 0, ck[base]);
    if (M > 1) {
      const double2 a1 = pc_col(k, nt, ns, 1, ay[base + int64_t(ky) * L]);
      const double2 a2 = pc_row(k, nt, ns, 1, ay[base + int64_t(ky ? M - ky : 0) * L]);
      acc.x += a1.x + a2.x;
      acc.y += a1.y + a2.y;
    }
    if (face) {
      const double2 f = face[nt];
      acc = make_double2(-(f.x * acc.x - f.y * acc.y), -(f.x * acc.y + f.y * acc.x));
      if (row == colq) acc.x += gram;
    }
    blk[e] = acc;
  }
}

// With ``face`` set, the entry becomes Gram*delta - face(n_t) * N2 entry
// (the face-averaged form inv(Gram*I - diag(face3)*circ_N) of
// Bragg_MOR_script.m:64-72, here with the exact axis last).
__global__ void k_pc_blocks(const double2* __restrict__ ck, double2* __restrict__ blk, int L, int M, int N,
                            int kxn, int kyn, const double2* __restrict__ twM, const double2* __restrict__ face,
                            double gram) {
  const int n3 = 3 * N;
  const int64_t total = int64_t(kxn) * kyn * n3 * n3;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    // output layout: block (kx, ky) at (kx*kyn + ky) * n3^2, column-major (row fastest)
    int64_t r = e;
    const int row = int(r % n3); r /= n3;
    const int colq = int(r % n3); r /= n3;
    const int ky = int(r % kyn); r /= kyn;
    const int kx = int(r);
    const int c = row / N, nt = row - c * N;
    const int d = colq / N, ns = colq - d * N;
    const int k = c_pc_comp[c][d];
    const int dz = nt > ns ? nt - ns : ns - nt;
    const double2* v = ck + (int64_t(k) * N + dz) * int64_t(M) * L + kx;   // v[t * L]
    double2 acc = pc_col(k, nt, ns, 0, v[0]);
    int64_t ph = ky;
    for (int t = 1; t < M; ++t, ph += ky) {
      if (ph >= M) ph -= M;
      const double2 ct = pc_col(k, nt, ns, t, v[int64_t(t) * L]);
      const double2 rt = pc_row(k, nt, ns, M - t, v[int64_t(M - t) * L]);
      const double wa = double(M - t) / M, wb = double(t) / M;
      const double2 cv = make_double2(wa * ct.x + wb * rt.x, wa * ct.y + wb * rt.y);
      const double2 w = twM[ph];
      acc.x += cv.x * w.x - cv.y * w.y;
      acc.y += cv.x * w.y + cv.y * w.x;
    }
    if (face) {
      const double2 f = face[nt];
      acc = make_double2(-(f.x * acc.x - f.y * acc.y), -(f.x * acc.y + f.y * acc.x));
      if (row == colq) acc.x += gram;
    }
    blk[e] = acc;
  }
}

// In-place Gauss-Jordan inversion with partial pivoting of one n x n
// column-major FP64 block per workgroup (in global memory).
__global__ void __launch_bounds__(256) k_pc_invert(double2* __restrict__ blk, int n, int* __restrict__ singular) {
  double2* A = blk + int64_t(blockIdx.x) * n * n;
  __shared__ int s_piv;
  __shared__ int s_perm[256];
  __shared__ double2 s_rowk[256];
  __shared__ double s_best[256];
  __shared__ int s_bi[256];
  for (int k = 0; k < n; ++k) {
    // pivot search over rows k..n-1 of column k
    double best = -1.0;
    int bi = k;
    for (int i = k + threadIdx.x; i < n; i += blockDim.x) {
      const double2 v = A[int64_t(k) * n + i];
      const double a = fabs(v.x) + fabs(v.y);
      if (a > best) { best = a; bi = i; }
    }
    s_best[threadIdx.x] = best;
    s_bi[threadIdx.x] = bi;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) {
        const double b2 = s_best[threadIdx.x + off];
        const int i2 = s_bi[threadIdx.x + off];
        if (b2 > s_best[threadIdx.x] || (b2 == s_best[threadIdx.x] && i2 < s_bi[threadIdx.x])) {
          s_best[threadIdx.x] = b2;
          s_bi[threadIdx.x] = i2;
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      s_piv = s_bi[0];
      s_perm[k] = s_bi[0];
      if (!(s_best[0] > 0.0)) singular[0] = 1;
    }
    __syncthreads();
    const int p = s_piv;
    // swap rows k and p
    if (p != k)
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const double2 t = A[int64_t(j) * n + k];
        A[int64_t(j) * n + k] = A[int64_t(j) * n + p];
        A[int64_t(j) * n + p] = t;
      }
    __syncthreads();
    const double2 piv = A[int64_t(k) * n + k];
    const double den = piv.x * piv.x + piv.y * piv.y;
    const double2 ip = den > 0 ? make_double2(piv.x / den, -piv.y / den) : make_double2(0.0, 0.0);
    __syncthreads();   // every wavefront has read the pivot before row k is overwritten
    // scale row k (pivot entry becomes 1/piv)
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      double2 v = (j == k) ? make_double2(1.0, 0.0) : A[int64_t(j) * n + k];
      v = cmul(v, ip);
      A[int64_t(j) * n + k] = v;
      s_rowk[j] = v;
    }
    __syncthreads();
    // eliminate column k from all other rows
    for (int64_t e = threadIdx.x; e < int64_t(n) * n; e += blockDim.x) {
      const int i = int(e % n);
      const int j = int(e / n);
      if (i == k) continue;
      const double2 f = A[int64_t(k) * n + i];
      if (j == k) {
        // new column k entry: -f * (1/piv)
        A[int64_t(j) * n + i] = make_double2(-(f.x * s_rowk[k].x - f.y * s_rowk[k].y),
                                             -(f.x * s_rowk[k].y + f.y * s_rowk[k].x));
      }
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < int64_t(n) * n; e += blockDim.x) {
      const int i = int(e % n);
      const int j = int(e / n);
      if (i == k || j == k) continue;
      // f was the old A[i][k]; recover it from the new column k entry: newAik = -f/piv  =>  f = -newAik*piv
      const double2 na = A[int64_t(k) * n + i];
      const double2 f = make_double2(-(na.x * piv.x - na.y * piv.y), -(na.x * piv.y + na.y * piv.x));
      const double2 rk = s_rowk[j];
      double2 v = A[int64_t(j) * n + i];
      v.x -= f.x * rk.x - f.y * rk.y;
      v.y -= f.x * rk.y + f.y * rk.x;
      A[int64_t(j) * n + i] = v;
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last first
  for (int k = n - 1; k >= 0; --k) {
    const int p = s_perm[k];
    if (p != k)
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double2 t = A[int64_t(k) * n + i];
        A[int64_t(k) * n + i] = A[int64_t(p) * n + i];
        A[int64_t(p) * n + i] = t;
      }
    __syncthreads();
  }
}

// 1-level blocks (level_1_parallel_func.m:5-169, level_1_parallel_func_N.m):
// for each kx the 3MN x 3MN matrix whose (component c, z index nt, y index m)
// x (d, ns, m') entry is the y-Toeplitz entry of component k = comp(c, d) at
// z offset |nt - ns|: the first column for m >= m' (t = m - m'), the first
// row for m < m' (t = m' - m), with the sign conventions of the 2-level blocks
// (pc_col, pc_row), v(t) = ck[k][|nt - ns|][t][kx].  Column-major per block.
// face == nullptr: the blocks of the folded tensor (mc, diag in ck);
// face = device complex[M*N] (index m + M n): Gram*I - diag(face3) * blocks
// of the bare tensor.
__global__ void k_pc_l1_blocks(const double2* __restrict__ ck, double2* __restrict__ blk, int L, int M, int N,
                               const double2* __restrict__ face, double gram) {
  const int MN = M * N, n3 = 3 * MN;
  const int64_t per = int64_t(n3) * n3;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < int64_t(L) * per;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int kx = int(e / per);
    const int64_t rc = e - int64_t(kx) * per;
    const int col = int(rc / n3), row = int(rc - int64_t(col) * n3);
    const int c = row / MN, rr = row - c * MN, nt = rr / M, m = rr - nt * M;
    const int d = col / MN, cr = col - d * MN, ns = cr / M, mp = cr - ns * M;
    const int k = c_pc_comp[c][d];
    const int dz = nt > ns ? nt - ns : ns - nt;
    const int t = m >= mp ? m - mp : mp - m;
    const double2 v = ck[((int64_t(k) * N + dz) * M + t) * L + kx];
    double2 a = m >= mp ? pc_col(k, nt, ns, t, v) : pc_row(k, nt, ns, t, v);
    if (face) {
      const double2 f = face[rr];
      a = make_double2(-(f.x * a.x - f.y * a.y), -(f.x * a.y + f.y * a.x));
      if (row == col) a.x += gram;
    }
    blk[e] = a;
  }
}

void build_level1_blocks(const double2* G, int L, int M, int N, double2 mc, double diag, double2* blk,
                         hipStream_t s, const double2* face_dev, double gram) {
  std::vector<double2> tL(L);
  for (int k = 0; k < L; ++k) {
    SinCos v = sincos_turn(k, L);
    tL[k] = make_double2(v.c, -v.s);
  }
  DevBuf dL(tL.size() * sizeof(double2));
  VP_HIP(hipMemcpyAsync(dL.p, tL.data(), dL.bytes, hipMemcpyHostToDevice, s));
  DevBuf ck(size_t(6) * L * M * N * sizeof(double2));
  const size_t lds2 = size_t(2) * L * sizeof(double2);
  if (lds2 <= 160 * 1024) {
    int L1 = 1;
    for (int d = 1; int64_t(d) * d <= L; ++d)
      if (L % d == 0) L1 = d;
    const void* fn = reinterpret_cast<const void*>(&k_pc_chan_x2);
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds2));
    hipLaunchKernelGGL(k_pc_chan_x2, dim3(6 * M * N), dim3(256), lds2, s, G, ck.as<double2>(), L, M, N, L1, mc, diag,
                       dL.as<double2>());
  } else {
    const int64_t total = int64_t(6) * L * M * N;
    hipLaunchKernelGGL(k_pc_chan_x, dim3(int(std::min<int64_t>((total + 255) / 256, 1 << 20))), dim3(256), 0, s, G,
                       ck.as<double2>(), L, M, N, mc, diag, dL.as<double2>());
  }
  VP_HIP(hipGetLastError());
  const int64_t n3 = 3 * int64_t(M) * N;
  const int64_t total = int64_t(L) * n3 * n3;
  hipLaunchKernelGGL(k_pc_l1_blocks, dim3(int(std::min<int64_t>((total + 255) / 256, 1 << 20))), dim3(256), 0, s,
                     ck.as<double2>(), blk, L, M, N, face_dev, gram);
  VP_HIP(hipGetLastError());
  VP_HIP(hipStreamSynchronize(s));
}

void build_level2_blocks(const double2* G, int L, int M, int N, double2 mc, double diag, double2* blk,
                         int kxn, int kyn, hipStream_t s, const double2* face_dev, double gram) {
  auto table = [](int P) {
    std::vector<double2> t(P);
    for (int k = 0; k < P; ++k) {
      SinCos v = sincos_turn(k, P);
      t[k] = make_double2(v.c, -v.s);
    }
    return t;
  };
  auto tL = table(L), tM = table(M);
  DevBuf dL(tL.size() * sizeof(double2)), dM(tM.size() * sizeof(double2));
  VP_HIP(hipMemcpyAsync(dL.p, tL.data(), dL.bytes, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dM.p, tM.data(), dM.bytes, hipMemcpyHostToDevice, s));
  DevBuf ck(size_t(6) * L * M * N * sizeof(double2));
  const char* ed = std::getenv("VP_PC_CHANX_DIRECT");   // 1: the direct O(L^2) kernel
  const size_t lds2 = size_t(2) * L * sizeof(double2);
  if (!(ed && std::atoi(ed) != 0) && lds2 <= 160 * 1024) {
    int L1 = 1;   // largest divisor of L not above sqrt(L)
    for (int d = 1; int64_t(d) * d <= L; ++d)
      if (L % d == 0) L1 = d;
    const void* fn = reinterpret_cast<const void*>(&k_pc_chan_x2);
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds2));
    hipLaunchKernelGGL(k_pc_chan_x2, dim3(6 * M * N), dim3(256), lds2, s, G, ck.as<double2>(), L, M, N, L1, mc, diag,
                       dL.as<double2>());
    VP_HIP(hipGetLastError());
  } else {
    const int64_t total = int64_t(6) * L * M * N;
    const int blocks = int(std::min<int64_t>((total + 255) / 256, 1 << 20));
    hipLaunchKernelGGL(k_pc_chan_x, dim3(blocks), dim3(256), 0, s, G, ck.as<double2>(), L, M, N, mc, diag,
                       dL.as<double2>());
    VP_HIP(hipGetLastError());
  }
  // y-Chan transform: direct sums (O(M) per entry) for short y axes, one FFT
  // per (component, dz, kx) line for long ones (VP_PC_CHANY=direct|fft|auto)
  int dev = 0;
  VP_HIP(hipGetDevice(&dev));
  const char* ey = std::getenv("VP_PC_CHANY");
  bool use_fft = M > 48;
  if (ey && !std::strcmp(ey, "direct")) use_fft = false;
  if (ey && !std::strcmp(ey, "fft")) use_fft = true;
  FftPlan yplan{};
  const size_t ldsy = size_t(2) * M * sizeof(double2);
  if (use_fft && (M < 2 || !factorize(M, true, yplan, 1) || ldsy > size_t(max_lds_bytes(dev)))) use_fft = false;
  const int64_t total = int64_t(kxn) * kyn * 9 * N * N;
  const int blocks = int(std::min<int64_t>((total + 255) / 256, 1 << 20));
  if (use_fft) {
    DevBuf ay(size_t(6) * L * M * N * sizeof(double2));
    const void* fy = reinterpret_cast<const void*>(&k_pc_chan_y_fft);
    (void)hipFuncSetAttribute(fy, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsy));
    hipLaunchKernelGGL(k_pc_chan_y_fft, dim3(unsigned(int64_t(6) * N * kxn)), dim3(256), ldsy, s, ck.as<double2>(),
                       ay.as<double2>(), L, M, kxn, yplan, dM.as<double2>());
    VP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pc_blocks_fft, dim3(blocks), dim3(256), 0, s, ck.as<double2>(), ay.as<double2>(), blk, L,
                       M, N, kxn, kyn, face_dev, gram);
    VP_HIP(hipGetLastError());
    VP_HIP(hipStreamSynchronize(s));
  } else {
    hipLaunchKernelGGL(k_pc_blocks, dim3(blocks), dim3(256), 0, s, ck.as<double2>(), blk, L, M, N, kxn, kyn,
                       dM.as<double2>(), face_dev, gram);
    VP_HIP(hipGetLastError());
  }
  VP_HIP(hipStreamSynchronize(s));
}

__global__ void k_parity_zero(const double2* __restrict__ G, int L, int M, int N, double* __restrict__ part) {
  const int64_t V = int64_t(L) * M * N;
  double zmax = 0.0, amax = 0.0;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < 6 * V;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int k = int(e / V);
    const int64_t v = e - k * V;
    const int l = int(v % L), m = int((v / L) % M), n = int(v / (int64_t(L) * M));
    const double a = fabs(G[e].x) + fabs(G[e].y);
    amax = fmax(amax, a);
    const bool on = (k == 1 && (l == 0 || m == 0)) || (k == 2 && (l == 0 || n == 0)) || (k == 4 && (m == 0 || n == 0));
    if (on) zmax = fmax(zmax, a);
  }
  __shared__ double sz[256], sa[256];
  sz[threadIdx.x] = zmax;
  sa[threadIdx.x] = amax;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) {
      sz[threadIdx.x] = fmax(sz[threadIdx.x], sz[threadIdx.x + o]);
      sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sz[0];
    part[2 * blockIdx.x + 1] = sa[0];
  }
}

bool toeplitz_parity_zero(const double2* G, int L, int M, int N, hipStream_t s) {
  const int64_t total = int64_t(6) * L * M * N;
  const int blocks = int(std::min<int64_t>((total + 255) / 256, 2048));
  DevBuf part(size_t(2) * blocks * sizeof(double));
  hipLaunchKernelGGL(k_parity_zero, dim3(blocks), dim3(256), 0, s, G, L, M, N, part.as<double>());
  VP_HIP(hipGetLastError());
  std::vector<double> h(size_t(2) * blocks);
  VP_HIP(hipMemcpyAsync(h.data(), part.p, part.bytes, hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  double z = 0.0, a = 0.0;
  for (int b = 0; b < blocks; ++b) {
    z = std::max(z, h[2 * b]);
    a = std::max(a, h[2 * b + 1]);
  }
  return std::isfinite(a) && z <= 1e-12 * a;
}

// Register-resident Gauss-Jordan inversion of one n x n block per workgroup
// (n <= 2*64 + 64 rows for the pivot search; the host keeps ceil(n/E)^2 <= NT).
// Thread (tr, tc) of a TR x TR grid, TR = ceil(n/E), holds rows tr + TR*a and
// columns tc + TR*b (a, b < E) in registers.  Pivoting is implicit: step k
// pivots on (p_k, k), p_k the unused row of largest |re|+|im| in column k
// (ties to the smaller row), and no rows move.  Each step is the exchange
//   W[p][k] = 1/piv, W[p][j] = W[p][j]/piv, W[i][k] = -W[i][k]/piv,
//   W[i][j] -= W[i][k] W[p][j]/piv,
// written as one rank-one update after zeroing row p and column k with
// c[p] = -1 and rs[k] = 1/piv.  After n steps inv(A)[pinv[i]][perm[j]] =
// W[i][j], with perm[k] = p_k and pinv its inverse.  Steps run column-major
// (k = r + TR*b, b unrolled), so the owners of column k read a compile-time
// register column; the pivot row index is wave uniform, so the owners of
// row p branch on it without divergence.
constexpr int kInvMaxN = 192;

template <int NT, int E>
__global__ void __launch_bounds__(NT) k_pc_invert_reg(double2* __restrict__ blk, int n, int* __restrict__ singular) {
  double2* A = blk + int64_t(blockIdx.x) * n * n;
  __shared__ double2 s_col[2][kInvMaxN];
  __shared__ double2 s_row[kInvMaxN];
  __shared__ int s_perm[kInvMaxN], s_pinv[kInvMaxN];
  const int TR = (n + E - 1) / E;
  const int tr = int(threadIdx.x) % TR, tc = int(threadIdx.x) / TR;
  const bool act = tc < TR;
  const int lane = int(threadIdx.x) & 63;
  double2 W[E][E];
#pragma unroll
  for (int a = 0; a < E; ++a)
#pragma unroll
    for (int b = 0; b < E; ++b) {
      const int i = tr + TR * a, j = tc + TR * b;
      W[a][b] = (act && i < n && j < n) ? A[int64_t(j) * n + i] : make_double2(0.0, 0.0);
    }
  unsigned used = 0;   // bit q: row lane + 64q already a pivot row
  int k = 0;
#pragma unroll
  for (int b = 0; b < E; ++b) {
    for (int r = 0; r < TR; ++r, ++k) {
      if (k >= n) break;
      const int buf = k & 1;
      // column k to LDS, then zero it (its new entries come from the update)
      if (act && tc == r) {
#pragma unroll
        for (int a = 0; a < E; ++a) {
          const int i = tr + TR * a;
          if (i < n) s_col[buf][i] = W[a][b];
          W[a][b] = make_double2(0.0, 0.0);
        }
      }
      __syncthreads();
      // every wavefront finds the same pivot
      double best = -1.0;
      int bi = 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int i = lane + 64 * q;
        if (i < n && !((used >> q) & 1u)) {
          const double2 v = s_col[buf][i];
          const double m = fabs(v.x) + fabs(v.y);
          if (m > best) { best = m; bi = i; }
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      const int p = __builtin_amdgcn_readfirstlane(bi);
      if ((p & 63) == lane) used |= 1u << (p >> 6);
      const double2 piv = s_col[buf][p];
      const double den = piv.x * piv.x + piv.y * piv.y;
      const bool ok = best > 0.0 && den > 0.0;
      const double2 ip = ok ? make_double2(piv.x / den, -piv.y / den) : make_double2(0.0, 0.0);
      if (threadIdx.x == 0) {
        s_perm[k] = p;
        s_pinv[p] = k;
        if (!ok) singular[0] = 1;
      }
      // owners of row p publish rs = row/piv (rs[k] = 1/piv) and zero the row
      const int pr = p % TR, ap = p / TR;
#pragma unroll
      for (int a = 0; a < E; ++a) {
        if (a == ap) {
          if (act && tr == pr) {
#pragma unroll
            for (int bb = 0; bb < E; ++bb) {
              const int j = tc + TR * bb;
              if (j < n) s_row[j] = (j == k) ? ip : cmul(W[a][bb], ip);
              W[a][bb] = make_double2(0.0, 0.0);
            }
          }
        }
      }
      __syncthreads();
      // rank-one update
      double2 cv[E], rv[E];
#pragma unroll
      for (int a = 0; a < E; ++a) {
        const int i = tr + TR * a;
        cv[a] = i < n ? s_col[buf][i] : make_double2(0.0, 0.0);
        if (i == p) cv[a] = make_double2(-1.0, 0.0);
      }
#pragma unroll
      for (int bb = 0; bb < E; ++bb) {
        const int j = tc + TR * bb;
        rv[bb] = (act && j < n) ? s_row[j] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int a = 0; a < E; ++a)
#pragma unroll
        for (int bb = 0; bb < E; ++bb) {
          W[a][bb].x -= cv[a].x * rv[bb].x - cv[a].y * rv[bb].y;
          W[a][bb].y -= cv[a].x * rv[bb].y + cv[a].y * rv[bb].x;
        }
    }
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int a = 0; a < E; ++a)
#pragma unroll
      for (int b = 0; b < E; ++b) {
        const int i = tr + TR * a, j = tc + TR * b;
        if (i < n && j < n) A[int64_t(s_perm[j]) * n + s_pinv[i]] = W[a][b];
      }
  }
}

// Blocked Gauss-Jordan for blocks too large for registers (n <= 192), one
// 256-thread workgroup per block, in place in global memory.  Columns are
// taken in panels of PB.  A panel (n x pb) is swept in LDS with the same
// implicit-pivot exchange as k_pc_invert_reg; that leaves D = [inv(W_PK);
// -W_OK inv(W_PK)] (P the panel's pivot rows, K its columns, O the others),
// and every other column j is then updated at once:
//   W[P][j] = D[P] W_old[P][j],  W[O][j] = W_old[O][j] + D[O] W_old[P][j],
// i.e. rows P cleared and the rank-pb product D * W_old[P][j] added.  Each
// column is read and written once per panel by one wavefront (lanes along
// rows, D in registers).  The result is inv(A) up to the pivot permutations,
// applied at the end by following the cycles of perm with two LDS columns:
// column perm[j] of the inverse is column j of W with row i moved to pinv[i].
template <int PB>
__global__ void __launch_bounds__(256, 2) k_pc_invert_panel(double2* __restrict__ blk, int n, int* __restrict__ singular) {
  constexpr int NT = 256, MAXN = 192, EP = (MAXN * PB + NT - 1) / NT;
  double2* A = blk + int64_t(blockIdx.x) * n * n;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* s_pan = reinterpret_cast<double2*>(smem_raw);   // [q][i], PB x n
  double2* s_prow = s_pan + PB * n;                          // [q][j], PB x n
  __shared__ int s_used[MAXN], s_perm[MAXN], s_pinv[MAXN], s_plist[PB];
  __shared__ int s_p;
  __shared__ double2 s_ip;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < n; i += NT) s_used[i] = 0;
  for (int c0 = 0; c0 < n; c0 += PB) {
    const int pb = min(PB, n - c0);
    for (int e = tid; e < pb * n; e += NT) s_pan[e] = A[int64_t(c0) * n + e];
    __syncthreads();
    for (int q = 0; q < pb; ++q) {
      const int k = c0 + q;
      if (wv == 0) {
        double best = -1.0;
        int bi = 0;
#pragma unroll
        for (int a = 0; a < MAXN / 64; ++a) {
          const int i = lane + 64 * a;
          if (i < n && !s_used[i]) {
            const double2 v = s_pan[q * n + i];
            const double m = fabs(v.x) + fabs(v.y);
            if (m > best) { best = m; bi = i; }
          }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          const double ob = __shfl_xor(best, off);
          const int oi = __shfl_xor(bi, off);
          if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (lane == 0) {
          const double2 piv = s_pan[q * n + bi];
          const double den = piv.x * piv.x + piv.y * piv.y;
          const bool ok = best > 0.0 && den > 0.0;
          s_ip = ok ? make_double2(piv.x / den, -piv.y / den) : make_double2(0.0, 0.0);
          if (!ok) singular[0] = 1;
          s_p = bi;
          s_used[bi] = 1;
          s_perm[k] = bi;
          s_pinv[bi] = k;
          s_plist[q] = bi;
        }
      }
      __syncthreads();
      const int p = s_p;
      const double2 ip = s_ip;
      double2 nv[EP];
#pragma unroll
      for (int r = 0; r < EP; ++r) {
        const int e = tid + NT * r;
        if (e < pb * n) {
          const int qq = e / n, i = e - qq * n;
          const double2 ci = (i == p) ? make_double2(-1.0, 0.0) : s_pan[q * n + i];
          const double2 rs = (qq == q) ? ip : cmul(s_pan[qq * n + p], ip);
          const double2 base = (i == p || qq == q) ? make_double2(0.0, 0.0) : s_pan[e];
          nv[r] = make_double2(base.x - (ci.x * rs.x - ci.y * rs.y), base.y - (ci.x * rs.y + ci.y * rs.x));
        }
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < EP; ++r) {
        const int e = tid + NT * r;
        if (e < pb * n) s_pan[e] = nv[r];
      }
      __syncthreads();
    }
    // old pivot rows outside the panel, then the panel's own columns back
    for (int e = tid; e < pb * n; e += NT) {
      const int q = e / n, j = e - q * n;
      if (j < c0 || j >= c0 + pb) s_prow[e] = A[int64_t(j) * n + s_plist[q]];
    }
    for (int e = tid; e < pb * n; e += NT) A[int64_t(c0) * n + e] = s_pan[e];
    __syncthreads();
    // rank-pb update of every other column, one column per wavefront at a time
    double2 d[MAXN / 64][PB];
    bool inP[MAXN / 64];
#pragma unroll
    for (int a = 0; a < MAXN / 64; ++a) {
      const int i = lane + 64 * a;
      inP[a] = i < n && s_used[i] && s_pinv[i] >= c0;
#pragma unroll
      for (int q = 0; q < PB; ++q) d[a][q] = (i < n && q < pb) ? s_pan[q * n + i] : make_double2(0.0, 0.0);
    }
    for (int j = wv; j < n; j += NT / 64) {
      if (j >= c0 && j < c0 + pb) continue;
      double2 acc[MAXN / 64];
#pragma unroll
      for (int a = 0; a < MAXN / 64; ++a) {
        const int i = lane + 64 * a;
        acc[a] = (i < n && !inP[a]) ? A[int64_t(j) * n + i] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int q = 0; q < PB; ++q) {
        if (q < pb) {
          const double2 pr = s_prow[q * n + j];   // same address in every lane
#pragma unroll
          for (int a = 0; a < MAXN / 64; ++a) {
            acc[a].x += d[a][q].x * pr.x - d[a][q].y * pr.y;
            acc[a].y += d[a][q].x * pr.y + d[a][q].y * pr.x;
          }
        }
      }
#pragma unroll
      for (int a = 0; a < MAXN / 64; ++a) {
        const int i = lane + 64 * a;
        if (i < n) A[int64_t(j) * n + i] = acc[a];
      }
    }
    __syncthreads();
  }
  // permutations: follow the cycles of perm over whole columns
  double2* bufA = s_pan;
  double2* bufB = s_prow;
  int* s_seen = s_used;   // reused: 0 = column not yet placed
  for (int i = tid; i < n; i += NT) s_seen[i] = 0;
  __syncthreads();
  for (int j0 = 0; j0 < n; ++j0) {
    if (s_seen[j0]) continue;   // uniform: read after a barrier
    for (int i = tid; i < n; i += NT) bufA[s_pinv[i]] = A[int64_t(j0) * n + i];
    __syncthreads();
    int cur = j0;
    do {
      const int nxt = s_perm[cur];
      if (nxt != j0)
        for (int i = tid; i < n; i += NT) bufB[s_pinv[i]] = A[int64_t(nxt) * n + i];
      __syncthreads();
      for (int i = tid; i < n; i += NT) A[int64_t(nxt) * n + i] = bufA[i];
      if (tid == 0) s_seen[nxt] = 1;
      __syncthreads();
      double2* t = bufA;
      bufA = bufB;
      bufB = t;
      cur = nxt;
    } while (cur != j0);
  }
}

constexpr int kInvPanel = 12;   // panel width of k_pc_invert_panel

// Kernel choice: E elements per thread side with ceil(n/E)^2 <= NT for the
// register kernel (1..6 with 256 threads, 7 = 6 with 512), 8 = panel kernel.
static int invert_reg_shape(int n, int& nt) {
  for (int e = 1; e <= 6; ++e) {
    const int tr = (n + e - 1) / e;
    if (tr * tr <= 256) { nt = 256; return e; }
  }
  if (n <= 132) { nt = 512; return 6; }   // ceil(132/6)^2 = 484
  if (n <= 192) return 8;                  // panel kernel
  return 0;
}

void invert_blocks(double2* blk, int64_t nblocks, int n, hipStream_t s) {
  VP_REQUIRE(n <= 256, ERR_UNSUPPORTED, "block inversion supports blocks up to 256 x 256");
  DevBuf flag(sizeof(int));
  VP_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
  int nt = 256;
  const char* ev = std::getenv("VP_PC_INVERT_REG");   // 0: the global-memory kernel for every size
  const int e = (ev && std::atoi(ev) == 0) ? 0 : invert_reg_shape(n, nt);
  for (int64_t b0 = 0; b0 < nblocks; b0 += 65535) {
    const int nb = int(std::min<int64_t>(65535, nblocks - b0));
    double2* p = blk + b0 * n * n;
    switch (nt == 512 && e == 6 ? 7 : e) {
      case 1: hipLaunchKernelGGL((k_pc_invert_reg<256, 1>), dim3(nb), dim3(256), 0, s, p, n, flag.as<int>()); break;
      case 2: hipLaunchKernelGGL((k_pc_invert_reg<256, 2>), dim3(nb), dim3(256), 0, s, p, n, flag.as<int>()); break;
      case 3: hipLaunchKernelGGL((k_pc_invert_reg<256, 3>), dim3(nb), dim3(256), 0, s, p, n, flag.as<int>()); break;
      case 4: hipLaunchKernelGGL((k_pc_invert_reg<256, 4>), dim3(nb), dim3(256), 0, s, p, n, flag.as<int>()); break;
      case 5: hipLaunchKernelGGL((k_pc_invert_reg<256, 5>), dim3(nb), dim3(256), 0, s, p, n, flag.as<int>()); break;
      case 6: hipLaunchKernelGGL((k_pc_invert_reg<256, 6>), dim3(nb), dim3(256), 0, s, p, n, flag.as<int>()); break;
      case 7: hipLaunchKernelGGL((k_pc_invert_reg<512, 6>), dim3(nb), dim3(512), 0, s, p, n, flag.as<int>()); break;
      case 8: {
        const void* fn = reinterpret_cast<const void*>(&k_pc_invert_panel<kInvPanel>);
        const int lds = 2 * kInvPanel * n * int(sizeof(double2));
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        hipLaunchKernelGGL((k_pc_invert_panel<kInvPanel>), dim3(nb), dim3(256), lds, s, p, n, flag.as<int>());
        break;
      }
      default: hipLaunchKernelGGL(k_pc_invert, dim3(nb), dim3(256), 0, s, p, n, flag.as<int>()); break;
    }
    VP_HIP(hipGetLastError());
  }
  int h = 0;
  VP_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  VP_REQUIRE(h == 0, ERR_INVALID, "preconditioner block is singular");
}

// rocSOLVER's strided batched complex LU and inverse, resolved from the
// shared libraries on first use (only 1-level blocks wider than 256 need
// them, so the library does not load rocBLAS otherwise).
namespace {
struct RocSolver {
  using create_t = rocblas_status (*)(rocblas_handle*);
  using destroy_t = rocblas_status (*)(rocblas_handle);
  using stream_t = rocblas_status (*)(rocblas_handle, hipStream_t);
  using getrf_t = decltype(&rocsolver_zgetrf_strided_batched);
  using getri_t = decltype(&rocsolver_zgetri_strided_batched);
  create_t create = nullptr;
  destroy_t destroy = nullptr;
  stream_t set_stream = nullptr;
  getrf_t getrf = nullptr;
  getri_t getri = nullptr;
  std::string error;
};

const RocSolver& rocsolver() {
  static const RocSolver r = [] {
    RocSolver o;
    void* hb = dlopen("librocblas.so", RTLD_NOW | RTLD_GLOBAL);
    if (!hb) hb = dlopen("/opt/rocm/lib/librocblas.so", RTLD_NOW | RTLD_GLOBAL);
    void* hs = dlopen("librocsolver.so", RTLD_NOW | RTLD_GLOBAL);
    if (!hs) hs = dlopen("/opt/rocm/lib/librocsolver.so", RTLD_NOW | RTLD_GLOBAL);
    if (!hb || !hs) {
      o.error = "rocBLAS / rocSOLVER libraries not found";
      return o;
    }
    o.create = reinterpret_cast<RocSolver::create_t>(dlsym(hb, "rocblas_create_handle"));
    o.destroy = reinterpret_cast<RocSolver::destroy_t>(dlsym(hb, "rocblas_destroy_handle"));
    o.set_stream = reinterpret_cast<RocSolver::stream_t>(dlsym(hb, "rocblas_set_stream"));
    o.getrf = reinterpret_cast<RocSolver::getrf_t>(dlsym(hs, "rocsolver_zgetrf_strided_batched"));
    o.getri = reinterpret_cast<RocSolver::getri_t>(dlsym(hs, "rocsolver_zgetri_strided_batched"));
    if (!o.create || !o.destroy || !o.set_stream || !o.getrf || !o.getri) o.error = "rocSOLVER symbols not found";
    return o;
  }();
  return r;
}
}  // namespace

void invert_blocks_any(double2* blk, int64_t nblocks, int n, hipStream_t s) {
  if (n <= 256) {
    invert_blocks(blk, nblocks, n, s);
    return;
  }
  const RocSolver& rs = rocsolver();
  VP_REQUIRE(rs.error.empty(), ERR_UNSUPPORTED, rs.error);
  VP_REQUIRE(nblocks <= INT32_MAX, ERR_UNSUPPORTED, "too many blocks");
  rocblas_handle h = nullptr;
  VP_REQUIRE(rs.create(&h) == rocblas_status_success, ERR_DEVICE, "rocblas_create_handle failed");
  int bad = 0;
  rocblas_status st1 = rocblas_status_success, st2 = rocblas_status_success;
  try {
    VP_REQUIRE(rs.set_stream(h, s) == rocblas_status_success, ERR_DEVICE, "rocblas_set_stream failed");
    DevBuf ipiv(size_t(nblocks) * n * sizeof(rocblas_int)), info(size_t(nblocks) * sizeof(rocblas_int));
    auto* A = reinterpret_cast<rocblas_double_complex*>(blk);
    const rocblas_stride sa = rocblas_stride(n) * n, sp = n;
    st1 = rs.getrf(h, n, n, A, n, sa, ipiv.as<rocblas_int>(), sp, info.as<rocblas_int>(), rocblas_int(nblocks));
    if (st1 == rocblas_status_success) {
      std::vector<rocblas_int> hi(nblocks);
      VP_HIP(hipMemcpyAsync(hi.data(), info.p, info.bytes, hipMemcpyDeviceToHost, s));
      VP_HIP(hipStreamSynchronize(s));
      for (rocblas_int v : hi) bad |= v != 0;
      if (!bad) {
        st2 = rs.getri(h, n, A, n, sa, ipiv.as<rocblas_int>(), sp, info.as<rocblas_int>(), rocblas_int(nblocks));
        VP_HIP(hipMemcpyAsync(hi.data(), info.p, info.bytes, hipMemcpyDeviceToHost, s));
        VP_HIP(hipStreamSynchronize(s));
        for (rocblas_int v : hi) bad |= v != 0;
      }
    }
  } catch (...) {
    rs.destroy(h);
    throw;
  }
  rs.destroy(h);
  VP_REQUIRE(st1 == rocblas_status_success && st2 == rocblas_status_success, ERR_DEVICE, "rocSOLVER call failed");
  VP_REQUIRE(!bad, ERR_INVALID, "preconditioner block is singular");
}

}  // namespace vp
