// JVIE N-operator matvec kernels: y = [Gram*Mr.*x - Mc.*N(x)](idx).
//
// Restates the arithmetic of pc/src_vie/src_mvp/mv_AN_const.m:102-132
// ('notransp') and :66-99 ('transp') as three fused passes over the device
// spectrum workspace X, laid out plane-major as X[kx][c][n][m]:
//   K1 k_xfwd  : gather x through the voxel map, zero pad in x, DFT along x
//                (two residue classes of length Lp/2), write X.
//   K2 k_plane : per kx plane, per (ky, kz) residue class: fold + twiddle,
//                2D DFT in LDS, 3x3 symmetric tensor multiply with fN,
//                inverse 2D DFT, unfold into register accumulators; the
//                cropped M x N plane is written back in place.
//   K3 k_xinv  : inverse DFT along x per class, crop to L, epilogue
//                Gram*Mr.*x - Mc.*N(x) and scatter to the unknown vector.
// fN is stored prescaled by 1/(Lp*Mp*Np), so no pass normalises.
#include "vp_internal.hpp"

namespace vp {

template <typename T>
__device__ __forceinline__ C<T> load_unknown(const ProbPtrs<T>& P, const C<T>* __restrict__ X, const Geometry& g,
                                             int64_t v, int c, int transp, int64_t& u_out) {
  int64_t u;
  int64_t pv;
  if (g.structured) {
    pv = P.pos[v];
    u = pv < 0 ? -1 : pv + int64_t(c) * g.nS;
  } else {
    pv = P.pos[int64_t(c) * g.V + v];
    u = pv;
  }
  u_out = u;
  if (u < 0) return czero<T>();
  C<T> xv = X[u];
  if (transp) xv = cmul(P.mc[pv], cconj(xv));  // fftn(conj(conj(Mc).*J)) = fftn(Mc.*conj(J))
  return xv;
}

// ----------------------------------------------------------------------------
// K1: forward x pass.  Block = B consecutive m-pencils at fixed (n, c).
// ----------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024) k_xfwd(const ProbPtrs<T>* __restrict__ probs, VecArgs<T> va, Geometry g,
                                               MatvecPlan mp, int transp, const C<T>* __restrict__ twx) {
  if (va.active && !va.active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  const C<T>* __restrict__ X = va.x + int64_t(blockIdx.y) * va.xstride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W = reinterpret_cast<C<T>*>(smem_raw);
  const int L = g.L, M = g.M, N = g.N;
  const int Ls = g.x.S, Lp = g.x.P, B = mp.B, ld = mp.ldsx;
  const int nmb = (M + B - 1) / B;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int mb = blk % nmb;
  const int rest = blk / nmb;
  const int n = rest % N;
  const int c = rest / N;
  const int m0 = mb * B;
  const int64_t MN = int64_t(M) * N;
  const LdsLayout lay{1, ld, 0, 1};
  C<T>* W1 = W + B * ld;
  for (int px = 0; px < g.x.R; ++px) {
    for (int e = threadIdx.x; e < B * Ls; e += blockDim.x) {
      const int bb = e / Ls;
      const int j = e - bb * Ls;
      const int m = m0 + bb;
      C<T> acc = czero<T>();
      if (m < M) {
        for (int l = j; l < L; l += Ls) {
          int64_t u;
          C<T> xv = load_unknown(P, X, g, g.voxel(l, m, n), c, transp, u);
          if (u >= 0) acc = cadd(acc, px ? cmul(xv, twx[(int64_t(l) * px) % Lp]) : xv);
        }
      }
      W[bb * ld + j] = acc;
    }
    __syncthreads();
    const C<T>* R = lds_fft<T, false>(W, W1, g.x.plan, B, lay, twx);
    for (int e = threadIdx.x; e < B * Ls; e += blockDim.x) {
      const int a = e / B;
      const int bb = e - a * B;
      const int m = m0 + bb;
      if (m < M) {
        const int64_t kxi = int64_t(px) * Ls + a;
        P.work[(kxi * 3 + c) * MN + int64_t(n) * M + m] = R[bb * ld + a];
      }
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// K2: one kx plane per workgroup.
// ----------------------------------------------------------------------------
// Launch bound for a kernel holding EPT complex accumulators per thread:
// large accumulators get 256 VGPRs (512 threads) instead of 128.
template <typename T, int EPT> constexpr int acc_bound() {
  return (sizeof(T) == 8 ? EPT <= 1 : EPT <= 4) ? 1024 : 512;
}

template <typename T, int EPT>
__global__ void __launch_bounds__((acc_bound<T, EPT>())) k_plane(const ProbPtrs<T>* __restrict__ probs, const int* __restrict__ active,
                                                Geometry g, const C<T>* __restrict__ twy,
                                                const C<T>* __restrict__ twz) {
  if (active && !active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W = reinterpret_cast<C<T>*>(smem_raw);
  const int M = g.M, N = g.N;
  const int Ms = g.y.S, Ns = g.z.S, Mp = g.y.P, Np = g.z.P, Ry = g.y.R, Rz = g.z.R;
  const int MN = M * N;
  const int MsNs = Ms * Ns;
  const int kxi = xcd_remap(blockIdx.x, gridDim.x);
  C<T>* Zp = P.work + int64_t(kxi) * 3 * MN;
  C<T> acc[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) acc[k] = czero<T>();
  const LdsLayout lay_y{1, Ms, 0, 1};
  const LdsLayout lay_z{Ms, MsNs, 1, Ms};
  C<T>* const W0 = W;
  C<T>* const W1 = W + 3 * MsNs;
  // twiddle tables live in LDS behind the two work buffers
  C<T>* const ty = W + 6 * MsNs;
  C<T>* const tz = ty + Mp;
  for (int e = threadIdx.x; e < Mp; e += blockDim.x) ty[e] = twy[e];
  for (int e = threadIdx.x; e < Np; e += blockDim.x) tz[e] = twz[e];
  const IDiv dMsNs(MsNs), dMs(Ms), dMN(MN), dM(M), dNs(Ns), dMp(Mp), dNp(Np);
  __syncthreads();
  for (int py = 0; py < Ry; ++py) {
    for (int pz = 0; pz < Rz; ++pz) {
      // fold + twiddle: W[c][z][y] = sum Z[c][y + i Ms][z + j Ns] w_Mp^{m py} w_Np^{n pz}
      for (int e = threadIdx.x; e < 3 * MsNs; e += blockDim.x) {
        const int c = dMsNs.div(e);
        const int r = e - c * MsNs;
        const int zz = dMs.div(r);
        const int yy = r - zz * Ms;
        C<T> v = czero<T>();
        for (int n = zz; n < N; n += Ns) {
          const C<T> wz = tz[dNp.mod(n * pz)];
          for (int m = yy; m < M; m += Ms) {
            C<T> t = Zp[(c * N + n) * M + m];
            if (py) t = cmul(t, ty[dMp.mod(m * py)]);
            if (pz) t = cmul(t, wz);
            v = cadd(v, t);
          }
        }
        W0[e] = v;
      }
      __syncthreads();
      C<T>* R = lds_fft<T, false>(W0, W1, g.y.plan, 3 * Ns, lay_y, ty);
      R = lds_fft<T, false>(R, R == W0 ? W1 : W0, g.z.plan, 3 * Ms, lay_z, tz);
      // 3x3 symmetric tensor multiply, components xx xy xz yy yz zz
      const C<T>* F = P.fN + ((int64_t(kxi) * Ry + py) * Rz + pz) * int64_t(MsNs) * 6;
      for (int e = threadIdx.x; e < MsNs; e += blockDim.x) {
        const C<T>* f = F + int64_t(e) * 6;
        const C<T> f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4], f5 = f[5];
        const C<T> a = R[e], b = R[MsNs + e], cc = R[2 * MsNs + e];
        R[e] = cadd(cadd(cmul(f0, a), cmul(f1, b)), cmul(f2, cc));
        R[MsNs + e] = cadd(cadd(cmul(f1, a), cmul(f3, b)), cmul(f4, cc));
        R[2 * MsNs + e] = cadd(cadd(cmul(f2, a), cmul(f4, b)), cmul(f5, cc));
      }
      __syncthreads();
      R = lds_fft<T, true>(R, R == W0 ? W1 : W0, g.z.plan, 3 * Ms, lay_z, tz);
      R = lds_fft<T, true>(R, R == W0 ? W1 : W0, g.y.plan, 3 * Ns, lay_y, ty);
      // unfold: out[c][n][m] += W[c][n % Ns][m % Ms] conj(w_Mp^{m py} w_Np^{n pz})
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int o = threadIdx.x + k * blockDim.x;
        if (o < 3 * MN) {
          const int c = dMN.div(o);
          const int r = o - c * MN;
          const int n = dM.div(r);
          const int m = r - n * M;
          C<T> t = R[(c * Ns + dNs.mod(n)) * Ms + dMs.mod(m)];
          if (py) t = cmulc(t, ty[dMp.mod(m * py)]);
          if (pz) t = cmulc(t, tz[dNp.mod(n * pz)]);
          acc[k] = cadd(acc[k], t);
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int o = threadIdx.x + k * blockDim.x;
    if (o < 3 * MN) Zp[o] = acc[k];
  }
}

// ----------------------------------------------------------------------------
// Split variant of the plane pass (y, z, y^-1 as separate sweeps).  The x
// spectra X[kx][c][n][m] are y-transformed into Y[kx][c][n][ky] (zero padded
// from M to Mp), the z pass works on tiles of Tz consecutive ky columns of one
// kx plane (zero pad N -> Np, DFT, tensor multiply, inverse, crop), and the
// inverse y sweep crops back to M into X.
// ----------------------------------------------------------------------------
// Generic y pass (runtime radix plan): By pencils per workgroup in two LDS
// buffers; the twiddles are read from the global table (cached), so a long
// line (Mp in the thousands) takes only its two buffers of LDS.
template <typename T, bool INV>
__global__ void __launch_bounds__(512) k_ypass(const ProbPtrs<T>* __restrict__ probs, const int* __restrict__ active,
                                               Geometry g, MatvecPlan mp, const C<T>* __restrict__ twy) {
  if (active && !active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);
  const int M = g.M, Mp = g.y.P, B = mp.By, ld = mp.ldy;
  C<T>* W1 = W0 + B * ld;
  const C<T>* tw = twy;
  const int64_t npen = int64_t(g.x.P) * 3 * g.N;
  const int64_t p0 = int64_t(xcd_remap(blockIdx.x, gridDim.x)) * B;
  const int nb = int(min<int64_t>(B, npen - p0));
  const IDiv dMp(Mp), dM(M);
  if (!INV) {
    for (int e = threadIdx.x; e < B * Mp; e += blockDim.x) {
      const int bb = dMp.div(e), k = e - bb * Mp;
      W0[bb * ld + k] = (bb < nb && k < M) ? P.work[(p0 + bb) * M + k] : czero<T>();
    }
  } else {
    for (int e = threadIdx.x; e < B * Mp; e += blockDim.x) {
      const int bb = dMp.div(e), k = e - bb * Mp;
      W0[bb * ld + k] = (bb < nb) ? P.work2[(p0 + bb) * Mp + k] : czero<T>();
    }
  }
  __syncthreads();
  const C<T>* R = lds_fft<T, INV>(W0, W1, g.y.plan, B, LdsLayout{1, ld, 0, 1}, tw);
  if (!INV) {
    for (int e = threadIdx.x; e < nb * Mp; e += blockDim.x) {
      const int bb = dMp.div(e), k = e - bb * Mp;
      P.work2[(p0 + bb) * Mp + k] = R[bb * ld + k];
    }
  } else {
    for (int e = threadIdx.x; e < nb * M; e += blockDim.x) {
      const int bb = dM.div(e), k = e - bb * M;
      P.work[(p0 + bb) * M + k] = R[bb * ld + k];
    }
  }
}

// Compile-time x passes over pencil-major x spectra X[c][n][m][kx], the
// class-px transform of pencil (c, n, m) at X[pencil][px*S + a] for
// kx = R*a + px (S = Lp/R = product of Rs, one pencil per workgroup, the
// transform contiguous in LDS).
// Forward: one workgroup per (pencil, px) gathers the L box voxels of the
// pencil through the voxel map, multiplies by w_Lp^{l px} (the first
// decimation-in-frequency step of the zero-padded length-Lp transform) and
// writes the S-point DFT.
template <typename T, int NT, int... Rs>
__global__ void __launch_bounds__(NT) k_xfwd_ct(const ProbPtrs<T>* __restrict__ probs, VecArgs<T> va, Geometry g,
                                                int transp, const C<T>* __restrict__ twx, int twlds) {
  constexpr int S = (Rs * ...);
  if (va.active && !va.active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  const C<T>* __restrict__ X = va.x + int64_t(blockIdx.y) * va.xstride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);     // one buffer: the stages run in place
  const int R = g.x.R, L = g.L, M = g.M, N = g.N;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int pencil = blk / R, px = blk - pencil * R;
  const int cn = pencil / M, m = pencil - cn * M;
  const int c = cn / N, n = cn - c * N;
  auto ld = [&](int, int l) {
    C<T> v = czero<T>();
    if (l < L) {
      int64_t u;
      v = load_unknown(P, X, g, g.voxel(l, m, n), c, transp, u);
      if (px) v = cmul(v, twx[l * px]);
    }
    return v;
  };
  const BufRes Wo = buf_res(P.work + int64_t(pencil) * g.x.P + int64_t(px) * S, uint32_t(S) * uint32_t(sizeof(C<T>)));
  auto st = [&](int, int k, C<T> v) { buf_store(Wo, uint32_t(k), v); };
  if (twlds) {
    // stage twiddles from two short FP64 factor tables in LDS instead of
    // the Lp-entry global table
    const TwSplit<T> tws = TwSplit<T>::fill(W0 + S, S * R);
    __syncthreads();
    ct_fft_ip<T, false, S, 1, 1, NT, 1, true, false, false, false, Rs...>(W0, tws, R, ld, st);
  } else {
    ct_fft_ip<T, false, S, 1, 1, NT, 1, true, false, false, false, Rs...>(W0, twx, R, ld, st);
  }
}

// x-inverse epilogue for box voxel (l, m, n), component c: the unknown's
// Gram*Mr.*x - Mc.*N(x) (or its 'transp' form, conjugated) from N(x) = nx.
template <typename T>
__device__ __forceinline__ void xinv_epilogue(const ProbPtrs<T>& P, const Geometry& g, const C<T>* __restrict__ X,
                                              C<T>* __restrict__ Y, int l, int m, int n, int c, int transp,
                                              C<T> nx) {
  const int64_t v = g.voxel(l, m, n);
  int64_t pv, u;
  if (g.structured) {
    pv = P.pos[v];
    u = pv < 0 ? -1 : pv + int64_t(c) * g.nS;
  } else {
    pv = P.pos[int64_t(c) * g.V + v];
    u = pv;
  }
  if (u >= 0) {
    const C<T> xv = X[u];
    C<T> id = cscale(xv, P.gram);
    if (P.mr) id = transp ? cmulc(id, P.mr[pv]) : cmul(id, P.mr[pv]);
    Y[u] = transp ? csub(id, cconj(nx)) : csub(id, cmul(P.mc[pv], nx));
  }
}

// Inverse: one workgroup per pencil runs the R class transforms, folds
// them back with conj(w_Lp^{l px}) in registers, crops to l < L and applies
// the epilogue Gram*Mr.*x - Mc.*N(x) (or its 'transp' form) per unknown.
template <typename T, int NT, int... Rs>
__global__ void __launch_bounds__(NT) k_xinv_ct(const ProbPtrs<T>* __restrict__ probs, VecArgs<T> va, Geometry g,
                                                int transp, const C<T>* __restrict__ twx, int twlds) {
  constexpr int S = (Rs * ...);
  constexpr int EPT = (S + NT - 1) / NT;
  if (va.active && !va.active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  const C<T>* __restrict__ X = va.x + int64_t(blockIdx.y) * va.xstride;
  C<T>* __restrict__ Y = va.y + int64_t(blockIdx.y) * va.ystride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);     // one buffer: the stages run in place
  const int R = g.x.R, L = g.L, M = g.M, N = g.N;
  const int pencil = xcd_remap(blockIdx.x, gridDim.x);
  const int cn = pencil / M, m = pencil - cn * M;
  const int c = cn / N, n = cn - c * N;
  C<T> acc[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = czero<T>();
  auto none = [](int, int, C<T>) {};
  TwSplit<T> tws{};
  if (twlds) {
    tws = TwSplit<T>::fill(W0 + S, S * R);
    __syncthreads();
  }
  for (int px = 0; px < R; ++px) {
    const BufRes Wi = buf_res(P.work + int64_t(pencil) * g.x.P + int64_t(px) * S, uint32_t(S) * uint32_t(sizeof(C<T>)));
    auto ld = [&](int, int k) { return buf_load<T>(Wi, uint32_t(k)); };
    const C<T>* Rb = twlds ? ct_fft_ip<T, true, S, 1, 1, NT, 1, true, false, false, true, Rs...>(W0, tws, R, ld, none)
                           : ct_fft_ip<T, true, S, 1, 1, NT, 1, true, false, false, true, Rs...>(W0, twx, R, ld, none);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int l = int(threadIdx.x) + e * NT;
      if (l < L) {
        C<T> t = Rb[l];
        if (px) t = cmulc(t, twx[l * px]);
        acc[e] = cadd(acc[e], t);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int l = int(threadIdx.x) + e * NT;
    if (l < L) xinv_epilogue(P, g, X, Y, l, m, n, c, transp, acc[e]);
  }
}

// Same for two residue classes (R = 2, FP32): both class transforms run as
// one batch of two in LDS (element k of class b at 2k + b), so a workgroup
// takes the stages once with twice the butterflies per stage.  The stages run
// in place in one buffer (2S elements), so two workgroups fit a CU.
template <typename T, int NT, int... Rs>
__global__ void __launch_bounds__(NT) k_xinv_ct2(const ProbPtrs<T>* __restrict__ probs, VecArgs<T> va, Geometry g,
                                                 int transp, const C<T>* __restrict__ twx) {
  constexpr int S = (Rs * ...);
  constexpr int EPT = (S + NT - 1) / NT;
  if (va.active && !va.active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  const C<T>* __restrict__ X = va.x + int64_t(blockIdx.y) * va.xstride;
  C<T>* __restrict__ Y = va.y + int64_t(blockIdx.y) * va.ystride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);     // one buffer: the stages run in place
  const int L = g.L, M = g.M, N = g.N;
  const int pencil = xcd_remap(blockIdx.x, gridDim.x);
  const int cn = pencil / M, m = pencil - cn * M;
  const int c = cn / N, n = cn - c * N;
  const TwSplit<T> tws = TwSplit<T>::fill(W0 + 2 * S, 2 * S);
  __syncthreads();
  const BufRes Wi = buf_res(P.work + int64_t(pencil) * g.x.P, uint32_t(2 * S) * uint32_t(sizeof(C<T>)));
  auto ld = [&](int b, int k) { return buf_load<T>(Wi, uint32_t(b * S + k)); };
  auto none = [](int, int, C<T>) {};
  const C<T>* Rb = ct_fft_ip<T, true, S, 2, 2, NT, 1, true, false, false, true, Rs...>(W0, tws, 2, ld, none);
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int l = int(threadIdx.x) + e * NT;
    if (l < L) {
      const C<T> t = cadd(Rb[2 * l], cmulc(Rb[2 * l + 1], twx[l]));
      xinv_epilogue(P, g, X, Y, l, m, n, c, transp, t);
    }
  }
}

// Compile-time y pass for Mp = product of Rs over NB pencils (kx, c, n):
// the box rows m < M of the x spectra and the full rows of the y spectra
// Y[kx][c][n][ky].  Forward: X -> Y (rows zero padded past M); inverse:
// Y -> X (cropped to M).  The x spectra sit either plane-major,
// X[kx][c][n][m] (XB false: a workgroup takes NB consecutive pencils), or
// pencil-major along x, X[c][n][m][kx] (XB true: NB consecutive kx of one
// (c, n), read and written b-fastest so a wavefront touches runs of NB
// consecutive kx).  Buffer resources cover the workgroup's data only, so a
// partial last group reads zeros and drops its stores.
template <typename T, bool INV, bool XB, int NB, int... Rs>
__global__ void __launch_bounds__(256) k_ypass_ct(const ProbPtrs<T>* __restrict__ probs,
                                                  const int* __restrict__ active, Geometry g,
                                                  const C<T>* __restrict__ twy) {
  constexpr int NLEN = (Rs * ...);
  constexpr int LD = NB + 1;
  constexpr uint32_t E = sizeof(C<T>);
  if (active && !active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);
  C<T>* W1 = W0 + NLEN * LD;
  C<T>* tw = W1 + NLEN * LD;
  for (int e = threadIdx.x; e < NLEN; e += 256) tw[e] = twy[e];
  const int M = g.M, N = g.N, Lp = g.x.P;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  // X element (b, m) at xo(b) + m*xs; Y element (b, ky) at b*ys + ky
  BufRes X, Y;
  uint32_t xs, ys;
  int nb;
  if constexpr (XB) {
    const int nkb = (Lp + NB - 1) / NB;
    const int cn = blk / nkb, kb = blk - cn * nkb;   // cn = c*N + n
    const int c = cn / N, n = cn - c * N;
    const int kx0 = kb * NB;
    nb = min(NB, Lp - kx0);
    X = buf_res(P.work + int64_t(cn) * M * Lp + kx0, uint32_t((M - 1) * Lp + nb) * E);
    xs = uint32_t(Lp);
    ys = uint32_t(3 * N * NLEN);
    Y = buf_res(P.work2 + (int64_t(kx0) * 3 * N + int64_t(c) * N + n) * NLEN, uint32_t((nb - 1) * ys + NLEN) * E);
  } else {
    const int64_t npen = int64_t(Lp) * 3 * N;
    const int64_t p0 = int64_t(blk) * NB;
    nb = int(min<int64_t>(NB, npen - p0));
    X = buf_res(P.work + p0 * M, uint32_t(nb) * uint32_t(M) * E);
    xs = 1;
    ys = uint32_t(NLEN);
    Y = buf_res(P.work2 + p0 * NLEN, uint32_t(nb) * uint32_t(NLEN) * E);
  }
  const uint32_t xb = XB ? 1u : uint32_t(M);   // X offset of pencil b: b * xb
  __syncthreads();
  if constexpr (!INV) {
    auto ld = [&](int b, int k) {
      return (k < M && b < nb) ? buf_load<T>(X, uint32_t(b) * xb + uint32_t(k) * xs) : czero<T>();
    };
    auto st = [&](int b, int k, C<T> v) { buf_store(Y, uint32_t(b) * ys + uint32_t(k), v); };
    ct_fft<T, false, NLEN, NB, LD, 256, 1, true, XB, false, false, Rs...>(W0, W1, tw, 1, ld, st);
  } else {
    auto ld = [&](int b, int k) { return buf_load<T>(Y, uint32_t(b) * ys + uint32_t(k)); };
    auto st = [&](int b, int k, C<T> v) {
      if (k < M && b < nb) buf_store(X, uint32_t(b) * xb + uint32_t(k) * xs, v);
    };
    ct_fft<T, true, NLEN, NB, LD, 256, 1, true, false, XB, false, Rs...>(W0, W1, tw, 1, ld, st);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_zpass(const ProbPtrs<T>* __restrict__ probs, const int* __restrict__ active,
                                               Geometry g, MatvecPlan mp, const C<T>* __restrict__ twz) {
  if (active && !active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W0 = reinterpret_cast<C<T>*>(smem_raw);
  const int N = g.N, Np = g.z.P, Mp = g.y.P, Tz = mp.Tz;
  const int nt = (Mp + Tz - 1) / Tz;
  const int Lp = g.x.P;
  // Logical order pairs kx with Lp - kx (0, 1, Lp-1, 2, Lp-2, ...): the two
  // planes read the same octant of a parity-compressed tensor, and with the
  // XCD remap their tiles run close together on one XCD (shared L2).  The
  // x spectra are stored by residue class: kx = Rx*a + px sits at px*Sx + a.
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int io = blk / nt;
  const int jo = (io + 1) >> 1;
  const int kx = io == 0 ? 0 : ((io & 1) ? jo : Lp - jo);
  const int kxi = (kx % g.x.R) * g.x.S + kx / g.x.R;
  const int ky0 = (blk - io * nt) * Tz;
  const int E = 3 * Tz * Np;
  C<T>* W1 = W0 + E;
  C<T>* tw = W1 + E;
  for (int e = threadIdx.x; e < Np; e += blockDim.x) tw[e] = twz[e];
  const IDiv dTz(Tz), dN(N);
  const int64_t base = int64_t(kxi) * 3 * N * Mp;   // Y[kx][c][n][ky]
  // LDS holds the 3*Tz columns batch-fastest: element n of column q = c*Tz + t
  // at n*ld + q, so loads, FFT stages and the multiply touch contiguous words.
  const int ld = 3 * Tz;
  for (int e = threadIdx.x; e < 3 * Tz * N; e += blockDim.x) {
    const int cn = dTz.div(e), t = e - cn * Tz;
    const int c = dN.div(cn), n = cn - c * N;
    const int ky = ky0 + t;
    W0[n * ld + c * Tz + t] = ky < Mp ? P.work2[base + (int64_t(c) * N + n) * Mp + ky] : czero<T>();
  }
  for (int e = threadIdx.x; e < ld * (Np - N); e += blockDim.x) W0[N * ld + e] = czero<T>();
  __syncthreads();
  C<T>* R = lds_fft_bf<T, false>(W0, W1, mp.zf, ld, ld, tw, N);
  // tensor multiply: fN is [kx][kz][ky][6] (class layout with Ry = Rz = 1) or,
  // parity compressed, the octant [kxc][kzc][comp][kyc] with the signs of the
  // odd components (xy: x,y; xz: x,z; yz: y,z) applied on mirrored axes.
  const bool mx = 2 * kx > Lp;
  const int kxc = mx ? Lp - kx : kx;
  const int Hy = Mp / 2 + 1, Hz = Np / 2 + 1;
  const C<T>* F = mp.fsym ? P.fN + int64_t(kxc) * Hz * Hy * 6 : P.fN + int64_t(kxi) * Np * Mp * 6;
  for (int e = threadIdx.x; e < Tz * Np; e += blockDim.x) {
    const int kz = dTz.div(e), t = e - kz * Tz;
    const int ky = ky0 + t;
    if (ky < Mp) {
      C<T> f0, f1, f2, f3, f4, f5;
      if (mp.fsym) {
        const bool my = 2 * ky > Mp, mz = 2 * kz > Np;
        const int kyc = my ? Mp - ky : ky, kzc = mz ? Np - kz : kz;
        const C<T>* f = F + int64_t(kzc) * 6 * Hy + kyc;     // component k at f[k * Hy]
        f0 = f[0]; f1 = f[Hy]; f2 = f[2 * Hy]; f3 = f[3 * Hy]; f4 = f[4 * Hy]; f5 = f[5 * Hy];
        if (mx != my) f1 = cneg(f1);
        if (mx != mz) f2 = cneg(f2);
        if (my != mz) f4 = cneg(f4);
      } else {
        const C<T>* f = F + (int64_t(kz) * Mp + ky) * 6;
        f0 = f[0]; f1 = f[1]; f2 = f[2]; f3 = f[3]; f4 = f[4]; f5 = f[5];
      }
      const int i0 = kz * ld + t, i1 = i0 + Tz, i2 = i0 + 2 * Tz;
      const C<T> a = R[i0], b = R[i1], cc = R[i2];
      R[i0] = cadd(cadd(cmul(f0, a), cmul(f1, b)), cmul(f2, cc));
      R[i1] = cadd(cadd(cmul(f1, a), cmul(f3, b)), cmul(f4, cc));
      R[i2] = cadd(cadd(cmul(f2, a), cmul(f4, b)), cmul(f5, cc));
    }
  }
  __syncthreads();
  const C<T>* Q = lds_fft_bf<T, true>(R, R == W0 ? W1 : W0, mp.zi, ld, ld, tw, 1 << 30, N);
  for (int e = threadIdx.x; e < 3 * Tz * N; e += blockDim.x) {
    const int cn = dTz.div(e), t = e - cn * Tz;
    const int c = dN.div(cn), n = cn - c * N;
    const int ky = ky0 + t;
    if (ky < Mp) P.work2[base + (int64_t(c) * N + n) * Mp + ky] = Q[n * ld + c * Tz + t];
  }
}

// ----------------------------------------------------------------------------
// Two-stage z pass for Np = RA * RB, fused around the tensor multiply.
// Forward plan [RA, RB] and inverse plan [RB, RA] (Stockham, natural order):
//   A : radix-RA butterflies j < RB on inputs n = j + RB*r (zero for n >= N),
//       read straight from the y spectra; outputs to LDS slots j*RA + k;
//   B : radix-RB butterflies j < RA on slots j + RA*r with twiddles w^{jr},
//       which leaves X[kz = j + RA*r] in registers for all three
//       components; the tensor multiply at those kz; then the first inverse
//       stage (radix RB, Ns = 1) runs on the same registers, outputs to
//       slots j*RB + k;
//   B': radix-RA butterflies j < RB on slots j + RB*r with twiddles w^{-jr};
//       outputs n = j + RB*r < N go straight back to the spectra.
// A workgroup takes TZ ky columns of one kx plane with tpk >= RA threads per
// column (thread = w*TZ + t): A and B' loop over the 3*RB (component,
// butterfly) items, B gives butterfly j = w to one thread for all three
// components.  One LDS buffer of Np x 3 x TZ values, element (slot, c, t) at
// slot*3*TZ + c*TZ + t; a wavefront's accesses are runs of TZ words.
// ----------------------------------------------------------------------------
template <typename T, int RA, int RB, int TZ>
__global__ void __launch_bounds__(256, (sizeof(T) == 8 ? 2 : 1)) k_zfused(const ProbPtrs<T>* __restrict__ probs, const int* __restrict__ active,
                                                Geometry g, MatvecPlan mp, const C<T>* __restrict__ twz) {
  constexpr int LD = 3 * TZ;
  if (active && !active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W = reinterpret_cast<C<T>*>(smem_raw);
  constexpr int TPK = zfused_threads(RA, RB);   // threads per ky column (launch width / TZ)
  const int N = g.N, Mp = g.y.P, Lp = g.x.P;
  const int nt = (Mp + TZ - 1) / TZ;
  // kx order and residue-class storage as in k_zpass
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int io = blk / nt;
  const int jo = (io + 1) >> 1;
  const int kx = io == 0 ? 0 : ((io & 1) ? jo : Lp - jo);
  const int kxi = (kx % g.x.R) * g.x.S + kx / g.x.R;
  const int t = threadIdx.x % TZ, w = threadIdx.x / TZ;
  const int ky = (blk - io * nt) * TZ + t;
  const bool kyok = ky < Mp;
  // the kx plane of the y spectra, Y[c][n][ky] at (c*N + n)*Mp + ky
  const uint32_t MpN = uint32_t(Mp) * N;
  const BufRes Y = buf_res(P.work2 + int64_t(kxi) * 3 * MpN, 3 * MpN * uint32_t(sizeof(C<T>)));
  const int rc = (N + RB - 1) / RB;                              // A inputs that can be nonzero

  // butterflies of A see at most rc nonzero inputs and B' keeps at most rc
  // outputs: pruned DFTs when that is at most half the radix
  const bool prune = 2 * rc <= RA;

  // ---- A: forward radix-RA stage from the spectra
  for (int it = w; it < 3 * RB; it += TPK) {
    const int c = it / RB, j = it - c * RB;
    C<T> v[RA];
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int n = j + RB * r;
      v[r] = czero<T>();
      if (r < rc && kyok && n < N) v[r] = buf_load<T>(Y, c * MpN + uint32_t(n) * Mp + ky);
    }
    if (prune) dft_in_pruned<T, RA, false>(v, rc);
    else dft<T, RA, false>(v);
#pragma unroll
    for (int k = 0; k < RA; ++k) W[(j * RA + k) * LD + c * TZ + t] = v[k];
  }
  __syncthreads();

  // ---- B: second forward stage, tensor multiply, first inverse stage
  C<T> x0[RB], x1[RB], x2[RB];
  const bool inb = w < RA;   // TPK >= RA: one stage-B butterfly per thread at most
  if (inb) {
    const int j = w;
    C<T> tw[RB];
#pragma unroll
    for (int r = 1; r < RB; ++r) tw[r] = twz[j * r];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int o = (j + RA * r) * LD + t;
      x0[r] = W[o];
      x1[r] = W[o + TZ];
      x2[r] = W[o + 2 * TZ];
      if (r > 0) {
        x0[r] = cmul(x0[r], tw[r]);
        x1[r] = cmul(x1[r], tw[r]);
        x2[r] = cmul(x2[r], tw[r]);
      }
    }
    dft<T, RB, false>(x0);
    dft<T, RB, false>(x1);
    dft<T, RB, false>(x2);
    if (kyok) {
      constexpr int Np = RA * RB;
      const bool mx = 2 * kx > Lp, my = 2 * ky > Mp;
      const int kxc = mx ? Lp - kx : kx, kyc = my ? Mp - ky : ky;
      const uint32_t Hy = Mp / 2 + 1, Hz = Np / 2 + 1;
      // tensor slab of this kx: octant [kzc][comp][kyc] or full [kz][ky][comp]
      const BufRes F = mp.fsym ? buf_res(P.fN + int64_t(kxc) * Hz * 6 * Hy, Hz * 6 * Hy * uint32_t(sizeof(C<T>)))
                               : buf_res(P.fN + int64_t(kxi) * Np * Mp * 6, uint32_t(Np) * Mp * 6 * sizeof(C<T>));
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int kz = j + RA * r;
        C<T> f0, f1, f2, f3, f4, f5;
        if (mp.fsym) {
          const bool mz = 2 * kz > Np;
          const uint32_t o = uint32_t(mz ? Np - kz : kz) * 6 * Hy + kyc;
          f0 = buf_load<T>(F, o); f1 = buf_load<T>(F, o + Hy); f2 = buf_load<T>(F, o + 2 * Hy);
          f3 = buf_load<T>(F, o + 3 * Hy); f4 = buf_load<T>(F, o + 4 * Hy); f5 = buf_load<T>(F, o + 5 * Hy);
          if (mx != my) f1 = cneg(f1);
          if (mx != mz) f2 = cneg(f2);
          if (my != mz) f4 = cneg(f4);
        } else {
          const uint32_t o = (uint32_t(kz) * Mp + ky) * 6;
          f0 = buf_load<T>(F, o); f1 = buf_load<T>(F, o + 1); f2 = buf_load<T>(F, o + 2);
          f3 = buf_load<T>(F, o + 3); f4 = buf_load<T>(F, o + 4); f5 = buf_load<T>(F, o + 5);
        }
        const C<T> a = x0[r], b = x1[r], cc = x2[r];
        x0[r] = cdot3(f0, a, f1, b, f2, cc);
        x1[r] = cdot3(f1, a, f3, b, f4, cc);
        x2[r] = cdot3(f2, a, f4, b, f5, cc);
        // double precision: keep at most a quarter of the tensor loads in
        // flight so the registers fit two waves per SIMD without spilling
        if constexpr (sizeof(T) == 8 && RB >= 8) {
          if ((r + 1) % (RB / 4) == 0) __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    dft<T, RB, true>(x0);
    dft<T, RB, true>(x1);
    dft<T, RB, true>(x2);
  }
  __syncthreads();   // every slot read before the buffer is reused
  if (inb) {
    const int j = w;
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const int o = (j * RB + k) * LD + t;
      W[o] = x0[k];
      W[o + TZ] = x1[k];
      W[o + 2 * TZ] = x2[k];
    }
  }
  __syncthreads();

  // ---- B': last inverse stage (radix RA), cropped to n < N
  for (int it = w; it < 3 * RB; it += TPK) {
    const int c = it / RB, j = it - c * RB;
    C<T> v[RA];
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      v[r] = W[(j + RB * r) * LD + c * TZ + t];
      if (r > 0) v[r] = cmul(v[r], twiddle<T, true>(twz, j * r));
    }
    if (prune) dft_out_pruned<T, RA, true>(v, rc);
    else dft<T, RA, true>(v);
    if (kyok) {
#pragma unroll
      for (int r = 0; r < RA; ++r) {
        const int n = j + RB * r;
        if (r < rc && n < N) buf_store(Y, c * MpN + uint32_t(n) * Mp + ky, v[r]);
      }
    }
  }
}

// ----------------------------------------------------------------------------
// Register-resident z pass for Np = R1 * R2.  One wavefront per component,
// TC = 64 / R1 ky columns per wavefront, column spread over R1 lanes: lane
// (l, t) holds x[R1*a + l], a < R2, of column ky0 + t.  Forward transform:
// R2-point DFT in registers, twiddle w^{l k1}, then a radix-2 decimation-in-
// frequency network across the R1 lanes (shuffles), which leaves
// X[k1 + R2*brev(l)] in lane l.  The three components meet in LDS for the
// tensor multiply; the inverse runs the mirror image (decimation-in-time
// network from bit-reversed order, conjugate twiddles, inverse R2 DFT) and
// stores rows n < N.  Twiddles come from the Np-entry table (LDS copy).
// ----------------------------------------------------------------------------
template <int R1> __device__ __forceinline__ int brev_lane(int l) {
  int r = 0;
#pragma unroll
  for (int b = 1; b < R1; b <<= 1) r = (r << 1) | ((l & b) ? 1 : 0);
  return r;
}

template <typename T> __device__ __forceinline__ C<T> shfl_xor_c(C<T> v, int mask) {
  C<T> o;
  o.x = __shfl_xor(v.x, mask);
  o.y = __shfl_xor(v.y, mask);
  return o;
}

// f_{c,d}(kx, ky, kz) for d = 0..2 (row c of the symmetric tensor).
template <typename T>
__device__ __forceinline__ void tensor_row(const ProbPtrs<T>& P, const MatvecPlan& mp, int c, int kxi, bool mx,
                                           int kxc, int ky, int kz, int Mp, int Np, C<T>& r0, C<T>& r1, C<T>& r2) {
  const C<T>* f;
  bool sxy = false, sxz = false, syz = false;
  int cs = 1;   // stride between components
  if (mp.fsym) {
    const bool my = 2 * ky > Mp, mz = 2 * kz > Np;
    const int Hy = Mp / 2 + 1, Hz = Np / 2 + 1;
    f = P.fN + (int64_t(kxc) * Hz + (mz ? Np - kz : kz)) * 6 * Hy + (my ? Mp - ky : ky);
    cs = Hy;
    sxy = mx != my;
    sxz = mx != mz;
    syz = my != mz;
  } else {
    f = P.fN + ((int64_t(kxi) * Np + kz) * Mp + ky) * 6;
  }
  if (c == 0) {
    r0 = f[0]; r1 = f[cs]; r2 = f[2 * cs];
    if (sxy) r1 = cneg(r1);
    if (sxz) r2 = cneg(r2);
  } else if (c == 1) {
    r0 = f[cs]; r1 = f[3 * cs]; r2 = f[4 * cs];
    if (sxy) r0 = cneg(r0);
    if (syz) r2 = cneg(r2);
  } else {
    r0 = f[2 * cs]; r1 = f[4 * cs]; r2 = f[5 * cs];
    if (sxz) r0 = cneg(r0);
    if (syz) r1 = cneg(r1);
  }
}

template <typename T, int R1, int R2>
__global__ void __launch_bounds__(192) k_zcol(const ProbPtrs<T>* __restrict__ probs, const int* __restrict__ active,
                                              Geometry g, MatvecPlan mp, const C<T>* __restrict__ twz) {
  constexpr int TC = 64 / R1;
  constexpr int Np = R1 * R2;
  constexpr bool PREFETCH = sizeof(T) == 4;   // FP32 keeps the tensor row in registers across the FFT
  if (active && !active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  __shared__ C<T> S[3][Np][TC];
  __shared__ C<T> tw[Np];
  for (int e = threadIdx.x; e < Np; e += blockDim.x) tw[e] = twz[e];
  const int N = g.N, Mp = g.y.P, Lp = g.x.P;
  const int nt = (Mp + TC - 1) / TC;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int io = blk / nt;
  const int jo = (io + 1) >> 1;
  const int kx = io == 0 ? 0 : ((io & 1) ? jo : Lp - jo);   // kx paired with Lp - kx (shared tensor octant)
  const int kxi = (kx % g.x.R) * g.x.S + kx / g.x.R;        // residue-class storage index of kx
  const bool mx = 2 * kx > Lp;
  const int kxc = mx ? Lp - kx : kx;
  const int ky0 = (blk - io * nt) * TC;
  const int c = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int l = lane / TC, t = lane - l * TC;
  const int ky = ky0 + t;
  const bool valid = ky < Mp;
  const int k2 = brev_lane<R1>(l);
  C<T>* Y = P.work2 + (int64_t(kxi) * 3 + c) * N * int64_t(Mp);
  C<T> fr[PREFETCH ? R2 : 1][3];
  if constexpr (PREFETCH) {
#pragma unroll
    for (int k1 = 0; k1 < R2; ++k1) {
      if (valid) tensor_row(P, mp, c, kxi, mx, kxc, ky, k1 + R2 * k2, Mp, Np, fr[k1][0], fr[k1][1], fr[k1][2]);
      else fr[k1][0] = fr[k1][1] = fr[k1][2] = czero<T>();
    }
  }
  C<T> v[R2];
#pragma unroll
  for (int a = 0; a < R2; ++a) {
    const int n = R1 * a + l;
    v[a] = (valid && n < N) ? Y[int64_t(n) * Mp + ky] : czero<T>();
  }
  __syncthreads();   // twiddle table
  dft<T, R2, false>(v);
#pragma unroll
  for (int k1 = 1; k1 < R2; ++k1) v[k1] = cmul(v[k1], tw[(l * k1) % Np]);
#pragma unroll
  for (int h = R1 / 2; h >= 1; h >>= 1) {
    const bool upper = (l & h) != 0;
    const C<T> w = tw[(l & (h - 1)) * (Np / (2 * h))];
#pragma unroll
    for (int k1 = 0; k1 < R2; ++k1) {
      const C<T> o = shfl_xor_c<T>(v[k1], h * TC);
      v[k1] = upper ? cmul(csub(o, v[k1]), w) : cadd(v[k1], o);
    }
  }
  // lane l holds X[k1 + R2 * k2]; components meet in LDS
#pragma unroll
  for (int k1 = 0; k1 < R2; ++k1) S[c][k1 + R2 * k2][t] = v[k1];
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R2; ++k1) {
    const int kz = k1 + R2 * k2;
    C<T> f0, f1, f2;
    if constexpr (PREFETCH) {
      f0 = fr[k1][0]; f1 = fr[k1][1]; f2 = fr[k1][2];
    } else {
      if (valid) tensor_row(P, mp, c, kxi, mx, kxc, ky, kz, Mp, Np, f0, f1, f2);
      else f0 = f1 = f2 = czero<T>();
    }
    v[k1] = cadd(cadd(cmul(f0, S[0][kz][t]), cmul(f1, S[1][kz][t])), cmul(f2, S[2][kz][t]));
  }
  // inverse: DIT network across lanes (bit-reversed in, natural out)
#pragma unroll
  for (int h = 1; h < R1; h <<= 1) {
    const bool upper = (l & h) != 0;
    const C<T> w = cconj(tw[(l & (h - 1)) * (Np / (2 * h))]);
#pragma unroll
    for (int k1 = 0; k1 < R2; ++k1) {
      const C<T> o = shfl_xor_c<T>(v[k1], h * TC);
      v[k1] = upper ? csub(o, cmul(w, v[k1])) : cadd(v[k1], cmul(w, o));
    }
  }
#pragma unroll
  for (int k1 = 1; k1 < R2; ++k1) v[k1] = cmulc(v[k1], tw[(l * k1) % Np]);
  dft<T, R2, true>(v);
#pragma unroll
  for (int a = 0; a < R2; ++a) {
    const int n = R1 * a + l;
    if (valid && n < N) Y[int64_t(n) * Mp + ky] = v[a];
  }
}

// ----------------------------------------------------------------------------
// K3: inverse x pass + epilogue.
// ----------------------------------------------------------------------------
template <typename T, int EPT>
__global__ void __launch_bounds__((acc_bound<T, EPT>())) k_xinv(const ProbPtrs<T>* __restrict__ probs, VecArgs<T> va, Geometry g,
                                               MatvecPlan mp, int transp, const C<T>* __restrict__ twx) {
  if (va.active && !va.active[blockIdx.y]) return;
  const ProbPtrs<T> P = probs[blockIdx.y];
  const C<T>* __restrict__ X = va.x + int64_t(blockIdx.y) * va.xstride;
  C<T>* __restrict__ Y = va.y + int64_t(blockIdx.y) * va.ystride;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C<T>* W = reinterpret_cast<C<T>*>(smem_raw);
  const int L = g.L, M = g.M, N = g.N;
  const int Ls = g.x.S, Lp = g.x.P, B = mp.B, ld = mp.ldsx;
  const int nmb = (M + B - 1) / B;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int mb = blk % nmb;
  const int rest = blk / nmb;
  const int n = rest % N;
  const int c = rest / N;
  const int m0 = mb * B;
  const int64_t MN = int64_t(M) * N;
  const LdsLayout lay{1, ld, 0, 1};
  C<T>* W1 = W + B * ld;
  C<T> acc[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) acc[k] = czero<T>();
  for (int px = 0; px < g.x.R; ++px) {
    for (int e = threadIdx.x; e < B * Ls; e += blockDim.x) {
      const int a = e / B;
      const int bb = e - a * B;
      const int m = m0 + bb;
      C<T> v = czero<T>();
      if (m < M) {
        const int64_t kxi = int64_t(px) * Ls + a;
        v = P.work[(kxi * 3 + c) * MN + int64_t(n) * M + m];
      }
      W[bb * ld + a] = v;
    }
    __syncthreads();
    const C<T>* R = lds_fft<T, true>(W, W1, g.x.plan, B, lay, twx);
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int o = threadIdx.x + k * blockDim.x;
      if (o < B * L) {
        const int bb = o / L;
        const int l = o - bb * L;
        C<T> t = R[bb * ld + l % Ls];
        if (px) t = cmulc(t, twx[(int64_t(l) * px) % Lp]);
        acc[k] = cadd(acc[k], t);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int o = threadIdx.x + k * blockDim.x;
    if (o < B * L) {
      const int bb = o / L;
      const int l = o - bb * L;
      const int m = m0 + bb;
      if (m < M) {
        const int64_t v = g.voxel(l, m, n);
        int64_t pv, u;
        if (g.structured) {
          pv = P.pos[v];
          u = pv < 0 ? -1 : pv + int64_t(c) * g.nS;
        } else {
          pv = P.pos[int64_t(c) * g.V + v];
          u = pv;
        }
        if (u >= 0) {
          const C<T> xv = X[u];
          C<T> id = cscale(xv, P.gram);
          if (P.mr) id = transp ? cmulc(id, P.mr[pv]) : cmul(id, P.mr[pv]);
          const C<T> nv = acc[k];
          Y[u] = transp ? csub(id, cconj(nv)) : csub(id, cmul(P.mc[pv], nv));
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------
// Launch
// ----------------------------------------------------------------------------
#define VP_EPT_DISPATCH(EPTV, CALL) \
  switch (EPTV) {                   \
    case 1: CALL(1); break;         \
    case 2: CALL(2); break;         \
    case 4: CALL(4); break;         \
    case 8: CALL(8); break;         \
    case 16: CALL(16); break;       \
    case 32: CALL(32); break;       \
    default: throw VpError{ERR_UNSUPPORTED, "matvec: accumulator size out of range"}; \
  }

template <typename T, bool INV>
static void launch_ypass_ct(const Geometry& g, const MatvecPlan& mp, const ProbPtrs<T>* d_probs, const int* active,
                            int nprob, const C<T>* twy, hipStream_t s) {
  constexpr int E = int(sizeof(C<T>));
  const int64_t npen = int64_t(g.x.P) * 3 * g.N;
  const bool xb = mp.xct != 0;
  switch (mp.yct) {
#define VP_YCT_LAUNCH(NL, ...)                                                                           \
  case NL: {                                                                                            \
    constexpr int NB = yct_pencils(NL, E);                                                              \
    const unsigned nblk = xb ? unsigned(3 * g.N * ((g.x.P + NB - 1) / NB)) : unsigned((npen + NB - 1) / NB); \
    if (xb) {                                                                                           \
      const void* fn = reinterpret_cast<const void*>(&k_ypass_ct<T, INV, true, NB, __VA_ARGS__>);       \
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, yct_lds(NL, E));        \
      hipLaunchKernelGGL((k_ypass_ct<T, INV, true, NB, __VA_ARGS__>), dim3(nblk, nprob), dim3(256),     \
                         yct_lds(NL, E), s, d_probs, active, g, twy);                                   \
    } else {                                                                                            \
      const void* fn = reinterpret_cast<const void*>(&k_ypass_ct<T, INV, false, NB, __VA_ARGS__>);      \
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, yct_lds(NL, E));        \
      hipLaunchKernelGGL((k_ypass_ct<T, INV, false, NB, __VA_ARGS__>), dim3(nblk, nprob), dim3(256),    \
                         yct_lds(NL, E), s, d_probs, active, g, twy);                                   \
    }                                                                                                   \
    break;                                                                                              \
  }
    VP_YCT_LIST(VP_YCT_LAUNCH)
#undef VP_YCT_LAUNCH
    default: throw VpError{ERR_UNSUPPORTED, "matvec: no compile-time y pass for this length"};
  }
}

// Compile-time x passes (class length mp.xct): forward one workgroup per
// (pencil, class), inverse one per pencil.
template <typename T, bool INV>
static void launch_xpass_ct(const Geometry& g, const MatvecPlan& mp, const ProbPtrs<T>* d_probs, VecArgs<T> va,
                            int nprob, int transp, const C<T>* twx, hipStream_t s) {
  const unsigned npen = unsigned(3 * g.N * g.M);
  const dim3 grid(INV ? npen : npen * unsigned(g.x.R), nprob);
  bool done = false;
  const int twlds = [] {   // VP_XTW_LDS=0: stage twiddles from the global table (read per launch)
    const char* e = std::getenv("VP_XTW_LDS");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  // FP32, two residue classes: both class transforms of a pencil in one batch
  const int xpair = [] {   // VP_XINV_PAIR=0: one class transform after the other (read per launch)
    const char* e = std::getenv("VP_XINV_PAIR");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  if constexpr (sizeof(T) == 4) {
    if (INV && xpair && g.x.R == 2) {
#define VP_XPAIR_LAUNCH(SL, NTP, PR)                                                                    \
  if (!done && mp.xct == SL &&                                                                          \
      2 * size_t(SL) * sizeof(C<T>) + TwSplit<T>::entries(2 * SL) * sizeof(double2) <= 160 * 1024) {    \
    const size_t lds = 2 * size_t(SL) * sizeof(C<T>) + TwSplit<T>::entries(2 * SL) * sizeof(double2);    \
    const void* fn = reinterpret_cast<const void*>(&k_xinv_ct2<T, NTP, VP_UNPAREN PR>);                 \
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));                \
    hipLaunchKernelGGL((k_xinv_ct2<T, NTP, VP_UNPAREN PR>), grid, dim3(NTP), lds, s, d_probs, va, g, transp, twx); \
    done = true;                                                                                        \
  }
      VP_XPAIR_LAUNCH(3600, 1024, (16, 9, 5, 5))
      VP_XPAIR_LAUNCH(1800, 768, (8, 9, 5, 5))
#undef VP_XPAIR_LAUNCH
    }
  }
#define VP_XCT_LAUNCH(SL, NTF, FR, NTI, IR)                                                             \
  if (!done && mp.xct == SL) {                                                                          \
    const size_t lds = size_t(SL) * sizeof(C<T>) + TwSplit<T>::entries(SL * g.x.R) * sizeof(double2);   \
    if (INV) {                                                                                          \
      const void* fn = reinterpret_cast<const void*>(&k_xinv_ct<T, NTI, VP_UNPAREN IR>);                \
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));              \
      hipLaunchKernelGGL((k_xinv_ct<T, NTI, VP_UNPAREN IR>), grid, dim3(NTI), lds, s, d_probs, va, g, transp, twx, \
                         twlds);                                                                        \
    } else {                                                                                            \
      const void* fn = reinterpret_cast<const void*>(&k_xfwd_ct<T, NTF, VP_UNPAREN FR>);                \
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));              \
      hipLaunchKernelGGL((k_xfwd_ct<T, NTF, VP_UNPAREN FR>), grid, dim3(NTF), lds, s, d_probs, va, g, transp, twx, \
                         twlds);                                                                        \
    }                                                                                                   \
    done = true;                                                                                        \
  }
  if constexpr (sizeof(T) == 8) {
    VP_XCT64_LIST(VP_XCT_LAUNCH)
  } else {
    VP_XCT32_LIST(VP_XCT_LAUNCH)
  }
#undef VP_XCT_LAUNCH
  if (!done) throw VpError{ERR_UNSUPPORTED, "matvec: no compile-time x pass for this length"};
}

template <typename T>
void launch_matvec(const Geometry& g, const MatvecPlan& mp, const ProbPtrs<T>* d_probs, VecArgs<T> va, int nprob,
                   int transp, const C<T>* twx, const C<T>* twy, const C<T>* twz, hipStream_t s) {
  const int nmb = (g.M + mp.B - 1) / mp.B;
  dim3 gx(nmb * g.N * 3, nprob);
  const bool xct = mp.variant == 1 && mp.xct != 0;
  if (xct) {
    ProfScope ps(s, KID_XFWD);
    launch_xpass_ct<T, false>(g, mp, d_probs, va, nprob, transp, twx, s);
    VP_HIP(hipGetLastError());
  } else {
    const void* fn = reinterpret_cast<const void*>(&k_xfwd<T>);
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(mp.lds1));
    ProfScope ps(s, KID_XFWD);
    hipLaunchKernelGGL(k_xfwd<T>, gx, dim3(mp.tx), mp.lds1, s, d_probs, va, g, mp, transp, twx);
    VP_HIP(hipGetLastError());
  }
  if (mp.variant == 1) {
    const int64_t npen = int64_t(g.x.P) * 3 * g.N;
    const dim3 gy(unsigned((npen + mp.By - 1) / mp.By), nprob);
    const void* fy0 = reinterpret_cast<const void*>(&k_ypass<T, false>);
    const void* fy1 = reinterpret_cast<const void*>(&k_ypass<T, true>);
    const void* fz = reinterpret_cast<const void*>(&k_zpass<T>);
    (void)hipFuncSetAttribute(fy0, hipFuncAttributeMaxDynamicSharedMemorySize, int(mp.lds_y));
    (void)hipFuncSetAttribute(fy1, hipFuncAttributeMaxDynamicSharedMemorySize, int(mp.lds_y));
    (void)hipFuncSetAttribute(fz, hipFuncAttributeMaxDynamicSharedMemorySize, int(mp.lds_z));
    {
      ProfScope ps(s, KID_YFWD);
      if (mp.yct) launch_ypass_ct<T, false>(g, mp, d_probs, va.active, nprob, twy, s);
      else hipLaunchKernelGGL((k_ypass<T, false>), gy, dim3(mp.ty), mp.lds_y, s, d_probs, va.active, g, mp, twy);
    }
    {
      ProfScope ps(s, KID_ZPASS);
      if (mp.zra > 0) {
        const dim3 gz(unsigned(g.x.P * ((g.y.P + kZfTile - 1) / kZfTile)), nprob);
        const dim3 bz(unsigned(kZfTile * mp.ztpk));
        bool zdone = false;
#define VP_ZF(A, B)                                                                                     \
  if (!zdone && mp.zra == A && mp.zrb == B) {                                                           \
    const void* fn = reinterpret_cast<const void*>(&k_zfused<T, A, B, kZfTile>);                        \
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(mp.lds_zf));          \
    hipLaunchKernelGGL((k_zfused<T, A, B, kZfTile>), gz, bz, mp.lds_zf, s, d_probs, va.active, g, mp, twz); \
    zdone = true;                                                                                       \
  }
        VP_ZF(11, 8) VP_ZF(11, 4) VP_ZF(11, 2) VP_ZF(7, 8) VP_ZF(7, 4) VP_ZF(7, 2) VP_ZF(5, 8) VP_ZF(5, 4)
        VP_ZF(5, 2) VP_ZF(3, 8) VP_ZF(3, 4) VP_ZF(3, 2) VP_ZF(3, 3) VP_ZF(8, 8) VP_ZF(8, 4) VP_ZF(4, 4)
        VP_ZF(4, 2)
        if constexpr (sizeof(T) == 4) {
          VP_ZF(11, 16) VP_ZF(7, 16) VP_ZF(5, 16) VP_ZF(3, 16) VP_ZF(16, 8)
        }
#undef VP_ZF
        if (!zdone) throw VpError{ERR_UNSUPPORTED, "matvec: no fused z kernel for this radix pair"};
      } else if (mp.zr1 > 0) {
        const dim3 gz(unsigned(g.x.P * ((g.y.P + 64 / mp.zr1 - 1) / (64 / mp.zr1))), nprob);
#define VP_ZCOL(A, B)                                                                                   \
  if (mp.zr1 == A && mp.zr2 == B) {                                                                     \
    hipLaunchKernelGGL((k_zcol<T, A, B>), gz, dim3(192), 0, s, d_probs, va.active, g, mp, twz);           \
    zdone = true;                                                                                       \
  }
        bool zdone = false;
        VP_ZCOL(8, 11) VP_ZCOL(8, 7) VP_ZCOL(8, 5) VP_ZCOL(8, 3) VP_ZCOL(8, 8) VP_ZCOL(8, 4)
        VP_ZCOL(4, 11) VP_ZCOL(4, 7) VP_ZCOL(4, 5) VP_ZCOL(4, 3) VP_ZCOL(4, 8) VP_ZCOL(4, 4)
        VP_ZCOL(16, 11) VP_ZCOL(16, 7) VP_ZCOL(16, 5) VP_ZCOL(16, 3) VP_ZCOL(16, 8) VP_ZCOL(16, 4)
#undef VP_ZCOL
        if (!zdone) throw VpError{ERR_UNSUPPORTED, "matvec: no register z-column kernel for this length"};
      } else {
        const dim3 gz(unsigned(g.x.P * ((g.y.P + mp.Tz - 1) / mp.Tz)), nprob);
        hipLaunchKernelGGL(k_zpass<T>, gz, dim3(mp.tz), mp.lds_z, s, d_probs, va.active, g, mp, twz);
      }
    }
    {
      ProfScope ps(s, KID_YINV);
      if (mp.yct) launch_ypass_ct<T, true>(g, mp, d_probs, va.active, nprob, twy, s);
      else hipLaunchKernelGGL((k_ypass<T, true>), gy, dim3(mp.ty), mp.lds_y, s, d_probs, va.active, g, mp, twy);
    }
    VP_HIP(hipGetLastError());
  } else {
    dim3 gp(g.x.P, nprob);
#define VP_K2(E)                                                                            \
  do {                                                                                      \
    const void* fn = reinterpret_cast<const void*>(&k_plane<T, E>);                         \
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(mp.lds2)); \
    ProfScope ps(s, KID_PLANE);                                                             \
    hipLaunchKernelGGL((k_plane<T, E>), gp, dim3(mp.t2), mp.lds2, s, d_probs, va.active, g, twy, twz);  \
  } while (0)
    VP_EPT_DISPATCH(mp.ept2, VP_K2);
#undef VP_K2
    VP_HIP(hipGetLastError());
  }
  if (xct) {
    ProfScope ps(s, KID_XINV);
    launch_xpass_ct<T, true>(g, mp, d_probs, va, nprob, transp, twx, s);
    VP_HIP(hipGetLastError());
  } else {
#define VP_K3(E)                                                                            \
  do {                                                                                      \
    const void* fn = reinterpret_cast<const void*>(&k_xinv<T, E>);                          \
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(mp.lds1)); \
    ProfScope ps(s, KID_XINV);                                                              \
    hipLaunchKernelGGL((k_xinv<T, E>), gx, dim3(mp.tx), mp.lds1, s, d_probs, va, g, mp, transp, twx); \
  } while (0)
    VP_EPT_DISPATCH(mp.ept3, VP_K3);
#undef VP_K3
    VP_HIP(hipGetLastError());
  }
}

template void launch_matvec<float>(const Geometry&, const MatvecPlan&, const ProbPtrs<float>*, VecArgs<float>, int,
                                   int, const C<float>*, const C<float>*, const C<float>*, hipStream_t);
template void launch_matvec<double>(const Geometry&, const MatvecPlan&, const ProbPtrs<double>*, VecArgs<double>,
                                    int, int, const C<double>*, const C<double>*, const C<double>*, hipStream_t);

}  // namespace vp
