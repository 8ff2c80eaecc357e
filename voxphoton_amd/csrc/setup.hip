// Operator setup on the device: geometry/plan selection, twiddle tables,
// circulant tensor layout, Toeplitz -> circulant spectrum, and assembly of the
// Toeplitz Green tensor.
#include <algorithm>
#include <memory>
#include <cstdlib>
#include <cmath>
#include <vector>
#include "op.hpp"

namespace vp {

// ----------------------------------------------------------------------------
// Geometry / plan
// ----------------------------------------------------------------------------
static std::vector<int> divisors(int n) {
  std::vector<int> d;
  for (int i = 1; i <= n; ++i)
    if (n % i == 0) d.push_back(i);
  return d;
}

void build_geometry(Geometry& g, MatvecPlan& mp, const vp_op_s& op, int Lp, int Mp, int Np, bool fp64, int device,
                    int variant) {
  g = Geometry{};
  g.L = op.Lb; g.M = op.Mb; g.N = op.Nb;
  g.l0 = op.l0; g.m0 = op.m0; g.n0 = op.n0;
  g.Lg = op.L; g.Mg = op.M;
  g.V = op.V;
  g.nS = op.nS;
  g.structured = op.structured;
  g.x.D = op.Lb; g.x.P = Lp;
  g.y.D = op.Mb; g.y.P = Mp;
  g.z.D = op.Nb; g.z.P = Np;
  VP_REQUIRE(Lp >= 2 * op.Lb - 1 && Mp >= 2 * op.Mb - 1 && Np >= 2 * op.Nb - 1, ERR_INVALID,
             "FFT sizes must be at least 2*B-1 for the unknowns' bounding box B");
  g.x.R = (Lp % 2 == 0) ? 2 : 1;
  g.x.S = Lp / g.x.R;
  VP_REQUIRE(g.x.S >= op.Lb, ERR_INVALID, "Lp too small for the unknowns' box");
  VP_REQUIRE(factorize(g.x.S, fp64, g.x.plan, g.x.R), ERR_UNSUPPORTED,
             "x FFT length " + std::to_string(g.x.S) + " has a prime factor above the supported radices");
  struct Cand { int ry, rz; };
  std::vector<Cand> cands;
  if (variant == 1) {
    // split passes: full padded lengths, no residue classes
    Geometry t = g;
    t.y.R = 1; t.y.S = Mp;
    t.z.R = 1; t.z.S = Np;
    std::string why2;
    VP_REQUIRE(factorize(t.y.S, fp64, t.y.plan, 1) && factorize(t.z.S, fp64, t.z.plan, 1), ERR_UNSUPPORTED,
               "padded y/z FFT lengths have a prime factor above the supported radices");
    MatvecPlan p{};
    VP_REQUIRE(make_matvec_plan(t, fp64, device, p, why2, 1), ERR_UNSUPPORTED, "matvec plan: " + why2);
    g = t;
    mp = p;
    return;
  }
  for (int ry : divisors(Mp))
    for (int rz : divisors(Np)) cands.push_back({ry, rz});
  std::stable_sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) {
    if (a.ry * a.rz != b.ry * b.rz) return a.ry * a.rz < b.ry * b.rz;
    return std::abs(a.ry - a.rz) < std::abs(b.ry - b.rz);
  });
  std::string why = "no residue-class split of the yz plane fits";
  for (const Cand& c : cands) {
    Geometry t = g;
    t.y.R = c.ry; t.y.S = Mp / c.ry;
    t.z.R = c.rz; t.z.S = Np / c.rz;
    if (!factorize(t.y.S, fp64, t.y.plan, t.y.R)) continue;
    if (!factorize(t.z.S, fp64, t.z.plan, t.z.R)) continue;
    MatvecPlan p{};
    if (!make_matvec_plan(t, fp64, device, p, why, 0)) continue;
    g = t;
    mp = p;
    return;
  }
  throw VpError{ERR_UNSUPPORTED, "matvec plan: " + why};
}

// exp(-2 pi i k / P) for k < P, correctly rounded to a few ulp.
static std::vector<double2> twiddle_table(int P) {
  std::vector<double2> t(P);
  for (int k = 0; k < P; ++k) {
    SinCos v = sincos_turn(k, P);
    t[k] = make_double2(v.c, -v.s);
  }
  return t;
}

template <typename T>
void upload_twiddles(OpPrec<T>& P, hipStream_t s) {
  const Geometry& g = P.g;
  std::vector<C<T>> h;
  h.reserve(g.x.P + g.y.P + g.z.P);
  for (int dimP : {g.x.P, g.y.P, g.z.P}) {
    auto t = twiddle_table(dimP);
    for (auto& v : t) h.push_back(mk<T>(T(v.x), T(v.y)));
  }
  P.tw.alloc(h.size() * sizeof(C<T>));
  VP_HIP(hipMemcpyAsync(P.tw.p, h.data(), h.size() * sizeof(C<T>), hipMemcpyHostToDevice, s));
  VP_HIP(hipStreamSynchronize(s));
}
template void upload_twiddles<float>(OpPrec<float>&, hipStream_t);
template void upload_twiddles<double>(OpPrec<double>&, hipStream_t);

// ----------------------------------------------------------------------------
// Class layout of the tensor
// ----------------------------------------------------------------------------
template <typename T>
__global__ void k_layout_fN(const double2* __restrict__ F, C<T>* __restrict__ out, Geometry g, double scale) {
  const int64_t total = int64_t(g.x.P) * g.y.P * g.z.P * 6;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = e;
    const int k = int(r % 6); r /= 6;
    const int yy = int(r % g.y.S); r /= g.y.S;
    const int zz = int(r % g.z.S); r /= g.z.S;
    const int pz = int(r % g.z.R); r /= g.z.R;
    const int py = int(r % g.y.R); r /= g.y.R;
    const int64_t kxi = r;
    const int px = int(kxi / g.x.S);
    const int a = int(kxi - int64_t(px) * g.x.S);
    const int64_t kx = int64_t(g.x.R) * a + px;
    const int64_t ky = int64_t(g.y.R) * yy + py;
    const int64_t kz = int64_t(g.z.R) * zz + pz;
    const double2 v = F[kx + g.x.P * (ky + int64_t(g.y.P) * (kz + int64_t(g.z.P) * k))];
    out[e] = mk<T>(T(v.x * scale), T(v.y * scale));
  }
}

template <typename T>
void layout_fN(const Geometry& g, const double2* F, DevBuf& out, double scale, hipStream_t s) {
  const int64_t total = int64_t(g.x.P) * g.y.P * g.z.P * 6;
  const int blocks = int(std::min<int64_t>((total + 255) / 256, 65536));
  if (out.bytes != size_t(total) * sizeof(C<T>)) out.alloc(size_t(total) * sizeof(C<T>));
  hipLaunchKernelGGL(k_layout_fN<T>, dim3(blocks), dim3(256), 0, s, F, out.as<C<T>>(), g, scale);
  VP_HIP(hipGetLastError());
}
template void layout_fN<double>(const Geometry&, const double2*, DevBuf&, double, hipStream_t);
template void layout_fN<float>(const Geometry&, const double2*, DevBuf&, double, hipStream_t);

// ----------------------------------------------------------------------------
// Toeplitz tensor -> circulant spectrum.  The circulant embedding of
// circulant_nop_const.m (reflected copies with the component parity, zero
// plane elsewhere) has, along each axis, the DFT
//   sum_i C(i) w^{ik} = G(0) + sum_{i>=1} G(i) (w^{ik} + s w^{-ik}),
// i.e. 2 cos for even and -2i sin for odd components; three separable
// passes evaluate it directly in FP64.
// ----------------------------------------------------------------------------
__constant__ int c_parity[6][3] = {{1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}, {1, -1, -1}, {1, 1, 1}};

// in : extents (D0, D1, D2) x 6 along the transformed axis "ax" (0,1,2), column-major
// out: same with extent P along ax.
__global__ void k_parity_dft(const double2* __restrict__ in, double2* __restrict__ out, int ax, int D, int P,
                             int e0, int e1, int e2, const double2* __restrict__ tw) {
  // output extents: e0,e1,e2 with e_ax == P ; input extents identical except e_ax == D
  const int64_t total = int64_t(e0) * e1 * e2 * 6;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = e;
    const int i0 = int(r % e0); r /= e0;
    const int i1 = int(r % e1); r /= e1;
    const int i2 = int(r % e2); r /= e2;
    const int k = int(r);
    int idx[3] = {i0, i1, i2};
    const int kk = idx[ax];
    int ein[3] = {e0, e1, e2};
    ein[ax] = D;
    int64_t stride = 1;
    for (int a = 0; a < ax; ++a) stride *= ein[a];
    idx[ax] = 0;
    const int64_t base = idx[0] + int64_t(ein[0]) * (idx[1] + int64_t(ein[1]) * (idx[2] + int64_t(ein[2]) * k));
    const bool odd = c_parity[k][ax] < 0;
    double2 acc = in[base];
    if (odd) acc = make_double2(0.0, 0.0);
    int64_t ph = kk;  // (i * kk) mod P, incremented
    for (int i = 1; i < D; ++i, ph += kk) {
      if (ph >= P) ph -= P;
      const double2 w = tw[ph];          // cos, -sin
      const double2 g = in[base + i * stride];
      if (!odd) {
        acc.x += 2.0 * w.x * g.x;
        acc.y += 2.0 * w.x * g.y;
      } else {
        // g * (-2 i sin) = g * (2 i w.y)
        acc.x += -2.0 * w.y * g.y;
        acc.y += 2.0 * w.y * g.x;
      }
    }
    out[e] = acc;
  }
}

// The x pass (the longest axis of the configurations) as an FFT: one
// workgroup per (m, n, component) pencil builds the embedded sequence C of
// length P = R*S (C(0) = G(0), zero for odd components; C(i) = G(i) and
// C(P-i) = s G(i) for 0 < i < D), folds it into R residue classes and runs
// an LDS FFT of length S per class: F(R a + p) = DFT_S(fold_p)(a).
__global__ void __launch_bounds__(256) k_embed_dft_x(const double2* __restrict__ G, double2* __restrict__ F, int D,
                                                     int P, int R, int S, FftPlan plan, int64_t npencil,
                                                     const double2* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* W0 = reinterpret_cast<double2*>(smem_raw);
  double2* W1 = W0 + S;
  const int64_t pen = blockIdx.x;
  if (pen >= npencil) return;
  const int k = int(pen / (npencil / 6));
  const bool odd = c_parity[k][0] < 0;
  const double sg = odd ? -1.0 : 1.0;
  const double2* g = G + pen * D;
  double2* f = F + pen * P;
  auto C = [&](int i) -> double2 {
    if (i == 0) return odd ? make_double2(0.0, 0.0) : g[0];
    if (i < D) return g[i];
    if (i > P - D) { const double2 v = g[P - i]; return make_double2(sg * v.x, sg * v.y); }
    return make_double2(0.0, 0.0);
  };
  const LdsLayout lay{1, S, 0, 1};
  for (int px = 0; px < R; ++px) {
    for (int j = threadIdx.x; j < S; j += blockDim.x) {
      double2 acc = make_double2(0.0, 0.0);
      for (int r = 0; r < R; ++r) {
        const int i = j + r * S;
        double2 v = C(i);
        if (px) v = cmul(v, tw[(int64_t(i) * px) % P]);
        acc = cadd(acc, v);
      }
      W0[j] = acc;
    }
    __syncthreads();
    const double2* Rr = lds_fft<double, false>(W0, W1, plan, 1, lay, tw);
    for (int a = threadIdx.x; a < S; a += blockDim.x) f[int64_t(R) * a + px] = Rr[a];
    __syncthreads();
  }
}

// The y and z passes as FFTs over strided lines: line (i, o) of the input
// (extent D along the axis, stride `inner`, base i + inner*D*o) is embedded
// as in k_embed_dft_x (C(0) = G(0), zero for components odd along the axis;
// C(i) = G(i), C(P-i) = s G(i)) and transformed at full length P in LDS;
// the output line has stride `inner` and base i + inner*P*o.  o runs over
// the outer extent times the six components (o / nouter6 is the component).
__global__ void __launch_bounds__(256) k_embed_dft_line(const double2* __restrict__ G, double2* __restrict__ F,
                                                        int ax, int D, int P, FftPlan plan, int64_t inner,
                                                        int64_t nouter6, const double2* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* W0 = reinterpret_cast<double2*>(smem_raw);
  double2* W1 = W0 + P;
  const int64_t line = blockIdx.x;
  const int64_t i = line % inner, o = line / inner;
  const int k = int(o / nouter6);
  const bool odd = c_parity[k][ax] < 0;
  const double sg = odd ? -1.0 : 1.0;
  const double2* g = G + i + inner * D * o;
  double2* f = F + i + inner * P * o;
  for (int j = threadIdx.x; j < P; j += blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (j == 0) {
      if (!odd) v = g[0];
    } else if (j < D) {
      v = g[int64_t(j) * inner];
    } else if (j > P - D) {
      const double2 u = g[int64_t(P - j) * inner];
      v = make_double2(sg * u.x, sg * u.y);
    }
    W0[j] = v;
  }
  __syncthreads();
  const LdsLayout lay{1, P, 0, 1};
  const double2* R = lds_fft<double, false>(W0, W1, plan, 1, lay, tw);
  for (int j = threadIdx.x; j < P; j += blockDim.x) f[int64_t(j) * inner] = R[j];
}

// Resampling of a circulant spectrum along one axis: line (i, o) of the
// input (length P0, stride inner, base i + inner*P0*o) is inverse
// transformed (the circulant's entries C(d), d < P0), the offsets
// -(B-1)..(B-1) are kept at their places in a circulant of length P1
// (negative offsets at P1 - |d|, zeros between), and the result transformed
// forward at length P1 (stride inner, base i + inner*P1*o).  Any circulant of
// at least 2B-1 points with the same offsets -(B-1)..(B-1) gives the same
// product on a box of extent B.
__global__ void __launch_bounds__(256) k_resample_line(const double2* __restrict__ in, double2* __restrict__ out,
                                                       int B, int P0, int P1, FftPlan pl0, FftPlan pl1,
                                                       int64_t inner, const double2* __restrict__ tw0,
                                                       const double2* __restrict__ tw1) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Pm = P0 > P1 ? P0 : P1;
  double2* W0 = reinterpret_cast<double2*>(smem_raw);
  double2* W1 = W0 + Pm;
  const int64_t line = blockIdx.x;
  const int64_t i = line % inner, o = line / inner;
  const double2* f = in + i + inner * P0 * o;
  double2* g = out + i + inner * P1 * o;
  for (int j = threadIdx.x; j < P0; j += blockDim.x) W0[j] = f[int64_t(j) * inner];
  __syncthreads();
  const LdsLayout lay0{1, P0, 0, 1};
  double2* R = lds_fft<double, true>(W0, W1, pl0, 1, lay0, tw0);
  double2* Q = R == W0 ? W1 : W0;
  const double inv = 1.0 / P0;
  for (int j = threadIdx.x; j < P1; j += blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (j < B) v = R[j];
    else if (j > P1 - B) v = R[P0 - (P1 - j)];
    Q[j] = make_double2(v.x * inv, v.y * inv);
  }
  __syncthreads();
  const LdsLayout lay1{1, P1, 0, 1};
  const double2* R2 = lds_fft<double, false>(Q, R, pl1, 1, lay1, tw1);
  for (int j = threadIdx.x; j < P1; j += blockDim.x) g[int64_t(j) * inner] = R2[j];
}

bool resample_fN(const double2* F, const int P0[3], const int B[3], const int P1[3], int device, DevBuf& out,
                 hipStream_t s) {
  FftPlan pl0[3]{}, pl1[3]{};
  const size_t lmax = size_t(max_lds_bytes(device));
  for (int a = 0; a < 3; ++a) {
    if (P0[a] == P1[a]) continue;
    if (P1[a] < 2 * B[a] - 1 || P0[a] < 2 * B[a] - 1) return false;
    if (!factorize(P0[a], true, pl0[a], 1) || !factorize(P1[a], true, pl1[a], 1) ||
        size_t(2) * std::max(P0[a], P1[a]) * sizeof(double2) > lmax)
      return false;
  }
  auto table = [&](int P) {
    auto t = twiddle_table(P);
    std::unique_ptr<DevBuf> d(new DevBuf(t.size() * sizeof(double2)));
    VP_HIP(hipMemcpyAsync(d->p, t.data(), d->bytes, hipMemcpyHostToDevice, s));
    VP_HIP(hipStreamSynchronize(s));   // t is freed on return
    return d;
  };
  const void* fk = reinterpret_cast<const void*>(&k_resample_line);
  int cur[3] = {P0[0], P0[1], P0[2]};
  const double2* in = F;
  DevBuf bufs[2];
  int nb = 0;
  for (int a = 0; a < 3; ++a) {
    if (P0[a] == P1[a]) continue;
    int nxt[3] = {cur[0], cur[1], cur[2]};
    nxt[a] = P1[a];
    DevBuf& o = bufs[nb++ & 1];
    o.release();
    o.alloc(size_t(nxt[0]) * nxt[1] * nxt[2] * 6 * sizeof(double2));
    auto tw0 = table(P0[a]), tw1 = table(P1[a]);
    // lines along axis a: stride = product of the extents before a; the
    // outer index runs over the extents after a and the 6 components
    const int64_t inner = a == 0 ? 1 : a == 1 ? int64_t(nxt[0]) : int64_t(nxt[0]) * nxt[1];
    const int64_t outer = a == 0 ? int64_t(nxt[1]) * nxt[2] * 6 : a == 1 ? int64_t(nxt[2]) * 6 : 6;
    const size_t lds = size_t(2) * std::max(P0[a], P1[a]) * sizeof(double2);
    (void)hipFuncSetAttribute(fk, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    hipLaunchKernelGGL(k_resample_line, dim3(unsigned(inner * outer)), dim3(256), lds, s, in, o.as<double2>(), B[a],
                       P0[a], P1[a], pl0[a], pl1[a], inner, tw0->as<double2>(), tw1->as<double2>());
    VP_HIP(hipGetLastError());
    VP_HIP(hipStreamSynchronize(s));   // the twiddle tables and the previous buffer are freed next
    in = o.as<double2>();
    cur[a] = P1[a];
  }
  if (nb == 0) return false;
  out = std::move(bufs[(nb - 1) & 1]);
  return true;
}

void toeplitz_to_fN(int L, int M, int N, int Lp, int Mp, int Np, int device, const double2* G, double2* F,
                    hipStream_t s) {
  VP_REQUIRE(Lp >= 2 * L - 1 && Mp >= 2 * M - 1 && Np >= 2 * N - 1, ERR_INVALID,
             "circulant sizes must be at least 2*D-1 for the tensor's extents D");
  auto twx = twiddle_table(Lp), twy = twiddle_table(Mp), twz = twiddle_table(Np);
  DevBuf dtx(twx.size() * sizeof(double2)), dty(twy.size() * sizeof(double2)), dtz(twz.size() * sizeof(double2));
  VP_HIP(hipMemcpyAsync(dtx.p, twx.data(), dtx.bytes, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dty.p, twy.data(), dty.bytes, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dtz.p, twz.data(), dtz.bytes, hipMemcpyHostToDevice, s));
  DevBuf t1(size_t(Lp) * M * N * 6 * sizeof(double2));
  DevBuf t2(size_t(Lp) * Mp * N * 6 * sizeof(double2));
  auto launch = [&](const double2* in, double2* out, int ax, int D, int P, int e0, int e1, int e2,
                    const double2* tw) {
    const int64_t total = int64_t(e0) * e1 * e2 * 6;
    const int blocks = int(std::min<int64_t>((total + 255) / 256, 1 << 20));
    hipLaunchKernelGGL(k_parity_dft, dim3(blocks), dim3(256), 0, s, in, out, ax, D, P, e0, e1, e2, tw);
    VP_HIP(hipGetLastError());
  };
  {
    // x pass by FFT when the residue-class length plans and fits LDS; the
    // direct sum otherwise
    const int R = (Lp % 2 == 0) ? 2 : 1;
    const int S = Lp / R;
    FftPlan plan{};
    const size_t lds = size_t(2) * S * sizeof(double2);
    if (factorize(S, true, plan, R) && lds <= size_t(max_lds_bytes(device)) && S >= L) {
      const int64_t npen = int64_t(M) * N * 6;
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_embed_dft_x),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
      hipLaunchKernelGGL(k_embed_dft_x, dim3(unsigned(npen)), dim3(256), lds, s, G, t1.as<double2>(), L, Lp, R, S,
                         plan, npen, dtx.as<double2>());
      VP_HIP(hipGetLastError());
    } else {
      launch(G, t1.as<double2>(), 0, L, Lp, Lp, M, N, dtx.as<double2>());
    }
  }
  // y and z passes: FFTs of the embedded lines where the length plans and
  // fits LDS and the direct sums would be long (D > 16), else direct sums
  auto pass = [&](const double2* in, double2* out, int ax, int D, int P, int e0, int e1, int e2, const double2* tw) {
    FftPlan plan{};
    const size_t lds = size_t(2) * P * sizeof(double2);
    if (D > 16 && factorize(P, true, plan, 1) && lds <= size_t(max_lds_bytes(device))) {
      const int64_t inner = ax == 1 ? int64_t(e0) : int64_t(e0) * e1;
      const int64_t nouter6 = ax == 1 ? int64_t(e2) : 1;
      const int64_t lines = inner * nouter6 * 6;
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_embed_dft_line),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
      hipLaunchKernelGGL(k_embed_dft_line, dim3(unsigned(lines)), dim3(256), lds, s, in, out, ax, D, P, plan, inner,
                         nouter6, tw);
      VP_HIP(hipGetLastError());
    } else {
      launch(in, out, ax, D, P, e0, e1, e2, tw);
    }
  };
  pass(t1.as<double2>(), t2.as<double2>(), 1, M, Mp, Lp, Mp, N, dty.as<double2>());
  pass(t2.as<double2>(), F, 2, N, Np, Lp, Mp, Np, dtz.as<double2>());
  VP_HIP(hipStreamSynchronize(s));
}

// ----------------------------------------------------------------------------
// Parity symmetry of the circulant spectrum.  The embedding of an even (odd)
// Toeplitz component is even (odd) along each axis, so
//   F_k(.., P - k_a, ..) = s_{k,a} F_k(.., k_a, ..)
// and the octant 0 <= k_a <= P_a/2 holds the whole tensor.  The check below
// measures max |F(e) - s F(mirror_a(e))| over all entries and axes against
// max |F|; a tensor given by the caller is compressed only if it passes.
// ----------------------------------------------------------------------------
__global__ void k_sym_check(const double2* __restrict__ F, int Lp, int Mp, int Np, double* __restrict__ part) {
  const int64_t total = int64_t(Lp) * Mp * Np * 6;
  double dmax = 0.0, amax = 0.0;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = e;
    const int kx = int(r % Lp); r /= Lp;
    const int ky = int(r % Mp); r /= Mp;
    const int kz = int(r % Np); r /= Np;
    const int k = int(r);
    const double2 v = F[e];
    amax = fmax(amax, fabs(v.x) + fabs(v.y));
    const int mir[3] = {kx ? Lp - kx : 0, ky ? Mp - ky : 0, kz ? Np - kz : 0};
    for (int a = 0; a < 3; ++a) {
      int q[3] = {kx, ky, kz};
      q[a] = mir[a];
      const double2 w = F[q[0] + int64_t(Lp) * (q[1] + int64_t(Mp) * (q[2] + int64_t(Np) * k))];
      const double sgn = c_parity[k][a];
      dmax = fmax(dmax, fabs(v.x - sgn * w.x) + fabs(v.y - sgn * w.y));
    }
  }
  __shared__ double sd[256], sa[256];
  sd[threadIdx.x] = dmax;
  sa[threadIdx.x] = amax;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) {
      sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + o]);
      sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd[0];
    part[2 * blockIdx.x + 1] = sa[0];
  }
}

bool fN_parity_symmetric(int Lp, int Mp, int Np, const double2* F, hipStream_t s) {
  const int64_t total = int64_t(Lp) * Mp * Np * 6;
  const int blocks = int(std::min<int64_t>((total + 255) / 256, 4096));
  DevBuf part(size_t(2) * blocks * sizeof(double));
  hipLaunchKernelGGL(k_sym_check, dim3(blocks), dim3(256), 0, s, F, Lp, Mp, Np, part.as<double>());
  VP_HIP(hipGetLastError());
  std::vector<double> h(size_t(2) * blocks);
  VP_HIP(hipMemcpyAsync(h.data(), part.p, part.bytes, hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  double dmax = 0.0, amax = 0.0;
  for (int b = 0; b < blocks; ++b) {
    dmax = std::max(dmax, h[2 * b]);
    amax = std::max(amax, h[2 * b + 1]);
  }
  return std::isfinite(dmax) && dmax <= 1e-12 * amax;
}

// Octant layout [kxc][kzc][comp][kyc], kxc <= Lp/2, kyc <= Mp/2, kzc <= Np/2:
// component-major inside a (kx, kz) slab, so work items on consecutive ky read
// consecutive words.
template <typename T>
__global__ void k_layout_fN_sym(const double2* __restrict__ F, C<T>* __restrict__ out, int Lp, int Mp, int Np,
                                double scale) {
  const int Hx = Lp / 2 + 1, Hy = Mp / 2 + 1, Hz = Np / 2 + 1;
  const int64_t total = int64_t(Hx) * Hy * Hz * 6;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    int64_t r = e;
    const int ky = int(r % Hy); r /= Hy;
    const int k = int(r % 6); r /= 6;
    const int kz = int(r % Hz); r /= Hz;
    const int kx = int(r);
    const double2 v = F[kx + int64_t(Lp) * (ky + int64_t(Mp) * (kz + int64_t(Np) * k))];
    out[e] = mk<T>(T(v.x * scale), T(v.y * scale));
  }
}

template <typename T>
void layout_fN_sym(const Geometry& g, const double2* F, DevBuf& out, double scale, hipStream_t s) {
  const int64_t total = int64_t(g.x.P / 2 + 1) * (g.y.P / 2 + 1) * (g.z.P / 2 + 1) * 6;
  const int blocks = int(std::min<int64_t>((total + 255) / 256, 65536));
  if (out.bytes != size_t(total) * sizeof(C<T>)) out.alloc(size_t(total) * sizeof(C<T>));
  hipLaunchKernelGGL(k_layout_fN_sym<T>, dim3(blocks), dim3(256), 0, s, F, out.as<C<T>>(), g.x.P, g.y.P, g.z.P,
                     scale);
  VP_HIP(hipGetLastError());
}
template void layout_fN_sym<double>(const Geometry&, const double2*, DevBuf&, double, hipStream_t);
template void layout_fN_sym<float>(const Geometry&, const double2*, DevBuf&, double, hipStream_t);

// ----------------------------------------------------------------------------
// Green tensor assembly.  For cell offset m (observation cell m, source cell
// 0) and faces p, q (ordered -x,+x,-y,+y,-z,+z), the dimensionless face-pair
// integral of e^{-j kappa rho}/rho is reduced through the difference
// variables of the two unit squares to a 2D (parallel faces) or 3D
// (perpendicular faces) integral with triangular weights Lam(s) = 1-|s|,
// split at the kinks of Lam so that a 1/rho singularity can only sit at a
// sub-box vertex, where a Duffy map removes it.
//   G_k(m) = dx^3/(4 pi) * sum_{p,q} I_pq(m) * P(p,q,k)
// with the projection P of vector_projection.m:7-22.
// ----------------------------------------------------------------------------
struct GaussRules {
  double x[4][16];
  double w[4][16];
  int n[4];   // 0 singular, 1 near, 2 medium, 3 far
};
__constant__ GaussRules c_rules;
__constant__ signed char c_proj[6][6][6];

__device__ __forceinline__ double lam(double s) { return 1.0 - fabs(s); }

// e^{-j kappa rho} / rho * jac for complex kappa = (re, im); im <= 0 damps.
__device__ __forceinline__ double2 kern(double2 kappa, double rho, double jac) {
  double sn, cs;
  sincos(kappa.x * rho, &sn, &cs);
  const double f = (kappa.y != 0.0 ? exp(kappa.y * rho) : 1.0) * jac / rho;
  return make_double2(cs * f, -sn * f);
}

// io >= 0 restricts the outermost quadrature index to io (a slice of the
// sum, for spreading one cell's integrals over threads); io < 0 sums all.
__device__ double2 pair_parallel(const int m[3], int a, int sp, int sq, double2 kappa, int rule, int io = -1) {
  const int b = (a == 0) ? 1 : 0;
  const int c = (a == 2) ? 1 : 2;
  const double h = double(sq - (m[a] + sp));
  const int mb = m[b], mc = m[c];
  double2 tot = make_double2(0.0, 0.0);
  for (int al = -1; al <= 0; ++al) {
    for (int be = -1; be <= 0; ++be) {
      const bool sing = (h == 0.0) && (mb == al || mb == al + 1) && (mc == be || mc == be + 1);
      if (sing) {
        const double v1 = mb, v2 = mc;
        const double sg1 = (mb == al) ? 1.0 : -1.0;
        const double sg2 = (mc == be) ? 1.0 : -1.0;
        const int n = c_rules.n[0];
        for (int i = io < 0 ? 0 : io, ie = io < 0 ? n : min(io + 1, n); i < ie; ++i) {
          const double xi = c_rules.x[0][i];
          for (int j = 0; j < n; ++j) {
            const double eta = c_rules.x[0][j];
            const double W = c_rules.w[0][i] * c_rules.w[0][j];
            for (int tri = 0; tri < 2; ++tri) {
              const double t1 = tri == 0 ? xi : xi * eta;
              const double t2 = tri == 0 ? xi * eta : xi;
              const double s1 = v1 + sg1 * t1, s2 = v2 + sg2 * t2;
              const double rho = sqrt(t1 * t1 + t2 * t2);
              const double2 g = kern(kappa, rho, xi);
              const double wl = W * lam(s1) * lam(s2);
              tot.x += wl * g.x;
              tot.y += wl * g.y;
            }
          }
        }
      } else {
        const int n = c_rules.n[rule];
        for (int i = io < 0 ? 0 : io, ie = io < 0 ? n : min(io + 1, n); i < ie; ++i) {
          const double s1 = al + c_rules.x[rule][i];
          const double d1 = s1 - mb;
          for (int j = 0; j < n; ++j) {
            const double s2 = be + c_rules.x[rule][j];
            const double d2 = s2 - mc;
            const double rho = sqrt(h * h + d1 * d1 + d2 * d2);
            const double2 g = kern(kappa, rho, 1.0);
            const double wl = c_rules.w[rule][i] * c_rules.w[rule][j] * lam(s1) * lam(s2);
            tot.x += wl * g.x;
            tot.y += wl * g.y;
          }
        }
      }
    }
  }
  return tot;
}

__device__ double2 pair_perp(const int m[3], int ap, int sp, int aq, int sq, double2 kappa, int rule, int io = -1) {
  const int c = 3 - ap - aq;
  const int ustar = m[ap] + sp;
  const int vstar = sq - m[aq];
  const int sstar = m[c];
  double2 tot = make_double2(0.0, 0.0);
  for (int ga = -1; ga <= 0; ++ga) {
    const bool sing = (ustar == 0 || ustar == 1) && (vstar == 0 || vstar == 1) && (sstar == ga || sstar == ga + 1);
    if (sing) {
      const double sgs = (sstar == ga) ? 1.0 : -1.0;
      const int n = c_rules.n[0];
      for (int i = io < 0 ? 0 : io, ie = io < 0 ? n : min(io + 1, n); i < ie; ++i) {
        const double xi = c_rules.x[0][i];
        for (int j = 0; j < n; ++j) {
          const double eta = c_rules.x[0][j];
          for (int k = 0; k < n; ++k) {
            const double zeta = c_rules.x[0][k];
            const double W = c_rules.w[0][i] * c_rules.w[0][j] * c_rules.w[0][k];
            for (int pyr = 0; pyr < 3; ++pyr) {
              double tu, tv, ts;
              if (pyr == 0) { tu = xi; tv = xi * eta; ts = xi * zeta; }
              else if (pyr == 1) { tu = xi * eta; tv = xi; ts = xi * zeta; }
              else { tu = xi * eta; tv = xi * zeta; ts = xi; }
              const double s = sstar + sgs * ts;
              const double rho = sqrt(tu * tu + tv * tv + ts * ts);
              const double2 g = kern(kappa, rho, xi * xi);
              const double wl = W * lam(s);
              tot.x += wl * g.x;
              tot.y += wl * g.y;
            }
          }
        }
      }
    } else {
      const int n = c_rules.n[rule];
      for (int i = io < 0 ? 0 : io, ie = io < 0 ? n : min(io + 1, n); i < ie; ++i) {
        const double d1 = c_rules.x[rule][i] - m[ap] - sp;
        for (int j = 0; j < n; ++j) {
          const double d2 = sq - m[aq] - c_rules.x[rule][j];
          for (int k = 0; k < n; ++k) {
            const double s = ga + c_rules.x[rule][k];
            const double d3 = s - m[c];
            const double rho = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
            const double2 g = kern(kappa, rho, 1.0);
            const double wl = c_rules.w[rule][i] * c_rules.w[rule][j] * c_rules.w[rule][k] * lam(s);
            tot.x += wl * g.x;
            tot.y += wl * g.y;
          }
        }
      }
    }
  }
  return tot;
}

// Cells with every offset below kNear use the singular and near rules; their
// integrals cost orders of magnitude more than a far cell's.
constexpr int kNear = 5;

// One workgroup per near cell: threads take (face p, face q, outermost
// quadrature index) items, and a fixed tree sums them.
__global__ void __launch_bounds__(256) k_assemble_near(double2* __restrict__ G, int L, int M, int N, double2 kappa,
                                                       double scale) {
  const int nl = min(L, kNear), nm = min(M, kNear);
  int b = blockIdx.x;
  const int i = b % nl;
  b /= nl;
  const int j = b % nm;
  const int l = b / nm;
  const int m[3] = {i, j, l};
  const int mx = max(i, max(j, l));
  const int rule = mx <= 1 ? 1 : 2;
  double acc[12];
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
  for (int it = threadIdx.x; it < 36 * 16; it += blockDim.x) {
    const int io = it % 16, pq = it / 16;
    const int p = pq / 6, q = pq - 6 * (pq / 6);
    const int ap = p >> 1, sp = p & 1, aq = q >> 1, sq = q & 1;
    bool any = false;
    for (int k = 0; k < 6; ++k) any |= c_proj[p][q][k] != 0;
    if (!any) continue;
    const double2 I = (ap == aq) ? pair_parallel(m, ap, sp, sq, kappa, rule, io)
                                 : pair_perp(m, ap, sp, aq, sq, kappa, rule, io);
    for (int k = 0; k < 6; ++k) {
      const double pr = c_proj[p][q][k];
      acc[2 * k] += pr * I.x;
      acc[2 * k + 1] += pr * I.y;
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 12; ++k) acc[k] += __shfl_xor(acc[k], off);
  __shared__ double part[4][12];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 12; ++k) part[wv][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double2 o = make_double2(0.0, 0.0);
    for (int w = 0; w < 4; ++w) { o.x += part[w][2 * k]; o.y += part[w][2 * k + 1]; }
    o = make_double2(o.x * scale, o.y * scale);
    bool zero = false;
    for (int ax = 0; ax < 3; ++ax) zero |= (c_parity[k][ax] < 0) && (m[ax] == 0);
    if (zero) o = make_double2(0.0, 0.0);
    const int64_t V = int64_t(L) * M * N;
    G[i + int64_t(L) * (j + int64_t(M) * l) + V * k] = o;
  }
}

// skip_near: cells left to k_assemble_near.
__global__ void k_assemble(double2* __restrict__ G, int L, int M, int N, double2 kappa, double scale, int skip_near) {
  const int64_t V = int64_t(L) * M * N;
  for (int64_t v = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; v < V; v += int64_t(gridDim.x) * blockDim.x) {
    const int i = int(v % L);
    const int j = int((v / L) % M);
    const int l = int(v / (int64_t(L) * M));
    const int m[3] = {i, j, l};
    const int mx = max(i, max(j, l));
    if (skip_near && mx < kNear) continue;
    const int rule = mx <= 1 ? 1 : (mx < kNear ? 2 : 3);
    double2 acc[6];
    for (int k = 0; k < 6; ++k) acc[k] = make_double2(0.0, 0.0);
    for (int p = 0; p < 6; ++p) {
      const int ap = p >> 1, sp = p & 1;
      for (int q = 0; q < 6; ++q) {
        const int aq = q >> 1, sq = q & 1;
        bool any = false;
        for (int k = 0; k < 6; ++k) any |= c_proj[p][q][k] != 0;
        if (!any) continue;
        const double2 I = (ap == aq) ? pair_parallel(m, ap, sp, sq, kappa, rule)
                                     : pair_perp(m, ap, sp, aq, sq, kappa, rule);
        for (int k = 0; k < 6; ++k) {
          const double pr = c_proj[p][q][k];
          acc[k].x += pr * I.x;
          acc[k].y += pr * I.y;
        }
      }
    }
    for (int k = 0; k < 6; ++k) {
      double2 o = make_double2(acc[k].x * scale, acc[k].y * scale);
      bool zero = false;
      for (int ax = 0; ax < 3; ++ax) zero |= (c_parity[k][ax] < 0) && (m[ax] == 0);
      if (zero) o = make_double2(0.0, 0.0);
      G[v + V * k] = o;
    }
  }
}

// Gauss-Legendre nodes and weights on [0, 1] (Newton iteration on the
// three-term recurrence of the Legendre polynomials).
static void gauss01(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-16) break;
    }
    // recompute derivative at converged root
    double p1 = 1.0, p2 = 0.0;
    for (int j = 1; j <= n; ++j) {
      double p3 = p2;
      p2 = p1;
      p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
    }
    pp = n * (z * p1 - p2) / (z * z - 1.0);
    // node z in (-1,1), descending; map to [0,1]
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * pp * pp);   // (2 / ((1-z^2) pp^2)) * 0.5
  }
}

static void init_assembly_constants() {
  static bool done[64] = {false};
  int dev = 0;
  VP_HIP(hipGetDevice(&dev));
  if (dev < 64 && done[dev]) return;
  GaussRules r{};
  const int orders[4] = {12, 10, 6, 3};
  for (int k = 0; k < 4; ++k) {
    r.n[k] = orders[k];
    gauss01(orders[k], r.x[k], r.w[k]);
  }
  VP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_rules), &r, sizeof(r)));
  signed char P[6][6][6];
  const int axes[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  for (int p = 0; p < 6; ++p)
    for (int q = 0; q < 6; ++q)
      for (int k = 0; k < 6; ++k) {
        // n_p x e_a . n_q x e_b ; n = sign * e_axis
        const int ap = p >> 1, aq = q >> 1;
        const int sgp = (p & 1) ? 1 : -1, sgq = (q & 1) ? 1 : -1;
        const int a = axes[k][0], b = axes[k][1];
        // (e_i x e_j) . (e_k x e_l) = d_ik d_jl - d_il d_jk
        int val = (ap == aq && a == b) - (ap == b && a == aq);
        P[p][q][k] = (signed char)(sgp * sgq * val);
      }
  VP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_proj), P, sizeof(P)));
  if (dev < 64) done[dev] = true;
}

void assemble_toeplitz_dev(int L, int M, int N, double dx, double ko, double2* G, hipStream_t s, double ko_im) {
  init_assembly_constants();
  const int64_t V = int64_t(L) * M * N;
  const int blocks = int(std::min<int64_t>((V + 127) / 128, 1 << 20));
  const double2 kappa = make_double2(ko * dx, ko_im * dx);
  const double scale = dx * dx * dx / (4.0 * kPi);
  const char* e = std::getenv("VP_ASSEMBLE_NEAR");   // 0: near cells integrated by one thread each
  const int near = (e && e[0] == '0') ? 0 : 1;
  if (near) {
    const int nnear = std::min(L, kNear) * std::min(M, kNear) * std::min(N, kNear);
    hipLaunchKernelGGL(k_assemble_near, dim3(nnear), dim3(256), 0, s, G, L, M, N, kappa, scale);
    VP_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_assemble, dim3(blocks), dim3(128), 0, s, G, L, M, N, kappa, scale, near);
  VP_HIP(hipGetLastError());
  VP_HIP(hipStreamSynchronize(s));
}

// ----------------------------------------------------------------------------
// Precision conversion
// ----------------------------------------------------------------------------
__global__ void k_c128_to_c64(const double2* __restrict__ in, float2* __restrict__ out, int64_t n, double scale) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    out[i] = make_float2(float(in[i].x * scale), float(in[i].y * scale));
}
__global__ void k_c64_to_c128(const float2* __restrict__ in, double2* __restrict__ out, int64_t n, double scale) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    out[i] = make_double2(double(in[i].x) * scale, double(in[i].y) * scale);
}
void convert_c128_to_c64(const double2* in, float2* out, int64_t n, double scale, hipStream_t s) {
  if (n <= 0) return;
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL(k_c128_to_c64, dim3(blocks), dim3(256), 0, s, in, out, n, scale);
  VP_HIP(hipGetLastError());
}
void convert_c64_to_c128(const float2* in, double2* out, int64_t n, double scale, hipStream_t s) {
  if (n <= 0) return;
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL(k_c64_to_c128, dim3(blocks), dim3(256), 0, s, in, out, n, scale);
  VP_HIP(hipGetLastError());
}

}  // namespace vp
