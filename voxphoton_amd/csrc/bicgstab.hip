// Preconditioned BiCGSTAB on the device: the 'B' solver of the reference's
// harness (JVIE_Solver_const.m:469-494, MATLAB's bicgstab(A,b,tol,maxit,M)).
// The iteration is the published one (H. A. van der Vorst, SIAM J. Sci.
// Stat. Comput. 13 (1992) 631-644, with the preconditioner applied to the
// search directions p and s):
//   rho = <rt, r>;  p = r + (rho/rho_old)(alpha/omega)(p - omega v)
//   ph = M^-1 p;  v = A ph;  alpha = rho / <rt, v>
//   s = r - alpha v;  x += alpha ph            (half step: |s| tested)
//   sh = M^-1 s;  t = A sh;  omega = <t, s> / <t, t>
//   x += omega sh;  r = s - omega t            (full step: |r| tested)
// with <a, b> = a^H b and the shadow residual rt = r0.  A candidate
// convergence of the recurrence is confirmed on the true residual b - A x;
// if it fails, the recurrence restarts from the true residual.
// Reductions are per-block partial sums added in a fixed order (no atomics).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <vector>
#include "krylov.hpp"
#include "capi_common.hpp"

namespace vp {
namespace {

constexpr int BT = 256;       // threads per block
constexpr int BMAX = 1024;    // partial-sum blocks

struct DotPairs {
  const void* a[4];
  const void* b[4];
  int npair;
};

// part[pair * nblk + block] = sum over the block's rows of conj(a) b
template <typename T>
__global__ void __launch_bounds__(BT) k_bcg_dots(DotPairs dp, int64_t n, double2* __restrict__ part) {
  double2 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = make_double2(0.0, 0.0);
  for (int64_t i = blockIdx.x * int64_t(BT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BT) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < dp.npair) {
        const C<T> a = static_cast<const C<T>*>(dp.a[q])[i];
        const C<T> b = static_cast<const C<T>*>(dp.b[q])[i];
        acc[q].x += double(a.x) * double(b.x) + double(a.y) * double(b.y);
        acc[q].y += double(a.x) * double(b.y) - double(a.y) * double(b.x);
      }
    }
  }
  __shared__ double2 sm[BT / 64][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    for (int off = 32; off > 0; off >>= 1) {
      acc[q].x += __shfl_xor(acc[q].x, off);
      acc[q].y += __shfl_xor(acc[q].y, off);
    }
  }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 4; ++q) sm[threadIdx.x >> 6][q] = acc[q];
  __syncthreads();
  if (threadIdx.x < 4 && int(threadIdx.x) < dp.npair) {
    double2 s = make_double2(0.0, 0.0);
    for (int w = 0; w < BT / 64; ++w) {
      s.x += sm[w][threadIdx.x].x;
      s.y += sm[w][threadIdx.x].y;
    }
    part[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] = s;
  }
}

__global__ void k_bcg_dots_final(const double2* __restrict__ part, int nblk, int npair, double2* __restrict__ out) {
  if (threadIdx.x < npair) {
    double2 s = make_double2(0.0, 0.0);
    for (int b = 0; b < nblk; ++b) {
      s.x += part[int64_t(threadIdx.x) * nblk + b].x;
      s.y += part[int64_t(threadIdx.x) * nblk + b].y;
    }
    out[threadIdx.x] = s;
  }
}

template <typename T>
__device__ __forceinline__ C<T> cmad(double2 a, C<T> x, C<T> y) {   // y + a x
  return mk<T>(T(double(y.x) + a.x * double(x.x) - a.y * double(x.y)),
               T(double(y.y) + a.x * double(x.y) + a.y * double(x.x)));
}

// p = r + beta (p - omega v)   (first iteration: p = r)
template <typename T>
__global__ void k_bcg_p(C<T>* __restrict__ p, const C<T>* __restrict__ r, const C<T>* __restrict__ v, double2 beta,
                        double2 omega, int first, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    if (first) {
      p[i] = r[i];
    } else {
      const C<T> d = cmad<T>(make_double2(-omega.x, -omega.y), v[i], p[i]);
      p[i] = cmad<T>(beta, d, r[i]);
    }
  }
}

// y = y + a x ; z = w - b u   (one pass over both updates)
template <typename T>
__global__ void k_bcg_step(C<T>* __restrict__ y, const C<T>* __restrict__ x, double2 a, C<T>* __restrict__ z,
                           const C<T>* __restrict__ w, const C<T>* __restrict__ u, double2 b, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    y[i] = cmad<T>(a, x[i], y[i]);
    z[i] = cmad<T>(make_double2(-b.x, -b.y), u[i], w[i]);
  }
}

template <typename T>
__global__ void k_bcg_sub(const C<T>* __restrict__ b, const C<T>* __restrict__ ax, C<T>* __restrict__ r, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    r[i] = mk<T>(b[i].x - ax[i].x, b[i].y - ax[i].y);
}

inline unsigned grid_n(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192))); }

inline std::complex<double> cd(double2 v) { return {v.x, v.y}; }
inline double2 d2(std::complex<double> v) { return make_double2(v.real(), v.imag()); }

}  // namespace

struct BicgResult {
  int flag = 1;
  int half_steps = 0;
  double relres = 0.0;
  std::vector<double> resvec;
};

template <typename T>
void bicgstab_device(ProblemSet<T>& ps, const C<T>* b, C<T>* x, bool x0_nonzero, double tol, int maxit,
                     BicgResult& res) {
  const int64_t n = ps.n;
  hipStream_t s = ps.stream;
  const size_t vb = size_t(n) * sizeof(C<T>);
  DevBuf r(vb), rt(vb), p(vb), v(vb), ph(vb), sv(vb), sh(vb), t(vb);
  const int nblk = int(std::min<int64_t>((n + BT - 1) / BT, BMAX));
  DevBuf part(size_t(4) * nblk * sizeof(double2)), dout(4 * sizeof(double2));
  auto dots = [&](std::initializer_list<std::pair<const C<T>*, const C<T>*>> pairs) {
    DotPairs dp{};
    dp.npair = 0;
    for (auto& pr : pairs) {
      dp.a[dp.npair] = pr.first;
      dp.b[dp.npair] = pr.second;
      ++dp.npair;
    }
    hipLaunchKernelGGL(k_bcg_dots<T>, dim3(nblk), dim3(BT), 0, s, dp, n, part.as<double2>());
    hipLaunchKernelGGL(k_bcg_dots_final, dim3(1), dim3(64), 0, s, part.as<double2>(), nblk, dp.npair,
                       dout.as<double2>());
    VP_HIP(hipGetLastError());
    std::vector<double2> h(dp.npair);
    VP_HIP(hipMemcpyAsync(h.data(), dout.p, dp.npair * sizeof(double2), hipMemcpyDeviceToHost, s));
    VP_HIP(hipStreamSynchronize(s));
    std::vector<std::complex<double>> out(dp.npair);
    for (int q = 0; q < dp.npair; ++q) out[q] = cd(h[q]);
    return out;
  };
  auto applyA = [&](const C<T>* in, C<T>* out) { ps_apply_A<T>(ps, in, n, out, n, nullptr); };
  auto applyM = [&](const C<T>* in, C<T>* out) {
    if (ps.has_prec) ps_apply_M<T>(ps, in, n, out, n, nullptr);
    else VP_HIP(hipMemcpyAsync(out, in, vb, hipMemcpyDeviceToDevice, s));
  };
  auto true_residual = [&]() {   // r = b - A x
    applyA(x, t.as<C<T>>());
    hipLaunchKernelGGL(k_bcg_sub<T>, dim3(grid_n(n)), dim3(256), 0, s, b, t.as<C<T>>(), r.as<C<T>>(), n);
    VP_HIP(hipGetLastError());
  };
  if (!x0_nonzero) VP_HIP(hipMemsetAsync(x, 0, vb, s));
  const double bnorm = std::sqrt(dots({{b, b}})[0].real());
  res.resvec.clear();
  if (!(bnorm > 0) || !std::isfinite(bnorm)) {
    // b = 0: x = 0 solves it; a non-finite b: unconverged (flag 1)
    if (bnorm == 0) VP_HIP(hipMemsetAsync(x, 0, vb, s));
    res.flag = bnorm == 0 ? 0 : 1;
    res.relres = bnorm == 0 ? 0.0 : relres_of(bnorm, bnorm);
    res.resvec.push_back(res.relres);
    VP_HIP(hipStreamSynchronize(s));
    return;
  }
  true_residual();
  VP_HIP(hipMemcpyAsync(rt.p, r.p, vb, hipMemcpyDeviceToDevice, s));
  auto d0 = dots({{r.as<C<T>>(), r.as<C<T>>()}});
  double rnorm = std::sqrt(d0[0].real());
  res.resvec.push_back(rnorm / bnorm);
  std::complex<double> rho_old = 1.0, alpha = 1.0, omega = 1.0, rho = d0[0];
  res.flag = 1;
  res.half_steps = 0;
  if (rnorm <= tol * bnorm) res.flag = 0;
  bool first = true;
  for (int it = 0; it < maxit && res.flag != 0; ++it) {
    if (!first) rho = dots({{rt.as<C<T>>(), r.as<C<T>>()}})[0];
    if (rho == 0.0 || !std::isfinite(std::abs(rho))) { res.flag = 4; break; }
    const std::complex<double> beta = first ? 0.0 : (rho / rho_old) * (alpha / omega);
    hipLaunchKernelGGL(k_bcg_p<T>, dim3(grid_n(n)), dim3(256), 0, s, p.as<C<T>>(), r.as<C<T>>(), v.as<C<T>>(),
                       d2(beta), d2(omega), first ? 1 : 0, n);
    VP_HIP(hipGetLastError());
    first = false;
    applyM(p.as<C<T>>(), ph.as<C<T>>());
    applyA(ph.as<C<T>>(), v.as<C<T>>());
    const std::complex<double> rtv = dots({{rt.as<C<T>>(), v.as<C<T>>()}})[0];
    if (rtv == 0.0 || !std::isfinite(std::abs(rtv))) { res.flag = 4; break; }
    alpha = rho / rtv;
    // x += alpha ph ; s = r - alpha v
    hipLaunchKernelGGL(k_bcg_step<T>, dim3(grid_n(n)), dim3(256), 0, s, x, ph.as<C<T>>(), d2(alpha),
                       sv.as<C<T>>(), r.as<C<T>>(), v.as<C<T>>(), d2(alpha), n);
    VP_HIP(hipGetLastError());
    const double snorm = std::sqrt(dots({{sv.as<C<T>>(), sv.as<C<T>>()}})[0].real());
    ++res.half_steps;
    res.resvec.push_back(snorm / bnorm);
    if (!std::isfinite(snorm)) { res.flag = 4; break; }
    if (snorm <= tol * bnorm) {
      true_residual();
      const double tn = std::sqrt(dots({{r.as<C<T>>(), r.as<C<T>>()}})[0].real());
      if (tn <= tol * bnorm) { res.flag = 0; break; }
      // the recurrence drifted: continue from the true residual
      rho = dots({{rt.as<C<T>>(), r.as<C<T>>()}})[0];
      first = true;
      continue;
    }
    applyM(sv.as<C<T>>(), sh.as<C<T>>());
    applyA(sh.as<C<T>>(), t.as<C<T>>());
    const auto ts = dots({{t.as<C<T>>(), sv.as<C<T>>()}, {t.as<C<T>>(), t.as<C<T>>()}});
    if (ts[1] == 0.0 || !std::isfinite(std::abs(ts[1]))) { res.flag = 4; break; }
    omega = ts[0] / ts[1];
    if (omega == 0.0) { res.flag = 4; break; }
    // x += omega sh ; r = s - omega t
    hipLaunchKernelGGL(k_bcg_step<T>, dim3(grid_n(n)), dim3(256), 0, s, x, sh.as<C<T>>(), d2(omega),
                       r.as<C<T>>(), sv.as<C<T>>(), t.as<C<T>>(), d2(omega), n);
    VP_HIP(hipGetLastError());
    const auto rr = dots({{r.as<C<T>>(), r.as<C<T>>()}, {rt.as<C<T>>(), r.as<C<T>>()}});
    rnorm = std::sqrt(rr[0].real());
    ++res.half_steps;
    res.resvec.push_back(rnorm / bnorm);
    if (!std::isfinite(rnorm)) { res.flag = 4; break; }
    rho_old = rho;
    rho = rr[1];
    if (rnorm <= tol * bnorm) {
      true_residual();
      const double tn = std::sqrt(dots({{r.as<C<T>>(), r.as<C<T>>()}})[0].real());
      if (tn <= tol * bnorm) { res.flag = 0; break; }
      rho = dots({{rt.as<C<T>>(), r.as<C<T>>()}})[0];
      first = true;
    }
  }
  // final unpreconditioned relative residual
  true_residual();
  const double tn = std::sqrt(dots({{r.as<C<T>>(), r.as<C<T>>()}})[0].real());
  res.relres = relres_of(tn, bnorm);
  if (res.flag == 0 && !(res.relres <= tol)) res.flag = 1;
  VP_HIP(hipStreamSynchronize(s));
}

}  // namespace vp

using namespace vp;

extern "C" {

int vp_bicgstab(vp_op op, vp_prec M, const void* b, const void* x0, void* x, int mem, int prec, double tol,
                int maxit, vp_gmres_info* info) {
  VP_API_BEGIN
  VP_REQUIRE(op && b && x, ERR_INVALID, "op/b/x is NULL");
  VP_REQUIRE(maxit > 0 && tol > 0, ERR_INVALID, "bad BiCGSTAB parameters");
  VP_REQUIRE(prec == VP_FP64 || prec == VP_FP32, ERR_INVALID, "prec must be VP_FP64 or VP_FP32");
  VP_REQUIRE(mem == VP_MEM_HOST || mem == VP_MEM_DEVICE, ERR_INVALID, "mem must be VP_MEM_HOST or VP_MEM_DEVICE");
  VP_HIP(hipSetDevice(op->ctx->device));
  hipStream_t s = op->ctx->stream;
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = op->n;
  vp_op_s* ops[1] = {op};
  vp_prec_s* precs[1] = {M};
  BicgResult res;
  auto run = [&](auto tag) {
    using T = decltype(tag);
    ProblemSet<T> ps;
    make_problem_set<T>(ps, ops, M ? precs : nullptr, 1);
    const double sc = std::is_same<T, float>::value ? 1.0 / op->scale32 : 1.0;   // b for A/scale
    if (mem == VP_MEM_DEVICE) {
      if (x0) VP_HIP(hipMemcpyAsync(x, x0, n * sizeof(C<T>), hipMemcpyDeviceToDevice, s));
      bicgstab_device<T>(ps, static_cast<const C<T>*>(b), static_cast<C<T>*>(x), x0 != nullptr, tol, maxit, res);
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
      bicgstab_device<T>(ps, db.as<C<T>>(), dx.as<C<T>>(), x0 != nullptr, tol, maxit, res);
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
  if (info) {
    info->flag = res.flag;
    info->relres = res.relres;
    info->total_iters = res.half_steps;
    info->iter_inner = res.half_steps / 2;
    info->iter_outer = res.half_steps % 2;
    info->ir_steps = 0;
    info->resvec_len = 0;
    if (info->resvec && info->resvec_cap > 0) {
      const int cnt = std::min<int>(info->resvec_cap, int(res.resvec.size()));
      for (int i = 0; i < cnt; ++i) info->resvec[i] = res.resvec[i];
      info->resvec_len = cnt;
    }
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  VP_API_END
}

}  // extern "C"
