// Field evaluation either side of the solve:
//   - the mollified dyadic Green's-function field of point currents (the
//     configs' complex-point-source beam), E_field_DGF_mollified.m:36-111;
//   - the total field Einc + (N(J) - Gram J) / (Gram j w eps0) through an
//     operator built over every voxel with unit contrast (mv_L_const.m:61-85
//     inside E_field_Nop_const.m:76-84).
// FP64 throughout; one thread per observation point.
#include <cmath>
#include <vector>
#include "op.hpp"
#include "capi_common.hpp"

namespace vp {
namespace {

constexpr double kMu0 = 4e-7 * 3.14159265358979323846;
constexpr double kC0 = 299792458.0;
constexpr double kEps0 = 1.0 / (kC0 * kC0 * kMu0);

__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double2 cexp(double2 z) {
  double s, c;
  sincos(z.y, &s, &c);
  const double e = exp(z.x);
  return make_double2(e * c, e * s);
}
// principal branch (non-negative real part)
__device__ __forceinline__ double2 csqrt(double2 z) {
  const double r = hypot(z.x, z.y);
  if (r == 0.0) return make_double2(0.0, 0.0);
  if (z.x >= 0.0) {
    const double t = sqrt(0.5 * (r + z.x));
    return make_double2(t, z.y / (2.0 * t));
  }
  const double t = sqrt(0.5 * (r - z.x));
  return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}

// Observation points: a voxel grid (origin + (l, m, n) dx, column-major
// voxel order) when oc is null, else the No x 3 column-major array oc.
struct ObsPoints {
  const double* oc;
  double ox, oy, oz, dx;
  int L, M;
};

__global__ void __launch_bounds__(256) k_dgf(const double2* __restrict__ J, const double2* __restrict__ ic, int Ni,
                                             ObsPoints obs, int64_t No, double ko, double omega, double cutoff,
                                             double2* __restrict__ E) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= No) return;
  double p[3];
  if (obs.oc) {
    p[0] = obs.oc[i];
    p[1] = obs.oc[No + i];
    p[2] = obs.oc[2 * No + i];
  } else {
    const int64_t lm = int64_t(obs.L) * obs.M;
    const int64_t n = i / lm, r = i - n * lm;
    const int64_t m = r / obs.L, l = r - m * obs.L;
    p[0] = obs.ox + double(l) * obs.dx;
    p[1] = obs.oy + double(m) * obs.dx;
    p[2] = obs.oz + double(n) * obs.dx;
  }
  double2 e[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  for (int s = 0; s < Ni; ++s) {
    double2 d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double2 q = ic[int64_t(c) * Ni + s];
      d[c] = make_double2(p[c] - q.x, -q.y);
    }
    const double2 R2 = cadd(cadd(cmul(d[0], d[0]), cmul(d[1], d[1])), cmul(d[2], d[2]));
    const double2 R = csqrt(R2);
    const double2 R3 = cmul(R, R2);
    const double2 psi =
        R.x < cutoff ? cexp(cneg(cdiv(make_double2(cutoff, 0.0), make_double2(cutoff - R2.x, -R2.y))))
                     : make_double2(0.0, 0.0);
    // chi = j w mu0 e^{-jkR} / (4 pi k^2 R^3)
    const double2 ejkr = cexp(make_double2(ko * R.y, -ko * R.x));
    const double2 chi = cdiv(cmul(make_double2(0.0, omega * kMu0), ejkr),
                             cscale(R3, 4.0 * 3.14159265358979323846 * ko * ko));
    const double2 P0 = make_double2(1.0 - ko * R.y, ko * R.x);           // j k R + 1
    const double2 Q0 = csub(cscale(R2, ko * ko), P0);                    // k^2 R2 - P0
    const double2 Pd = cdiv(csub(Q0, cscale(P0, 2.0)), R2);              // (Q0 - 2 P0) / R2
    const double2 Q = cmul(cmul(chi, Q0), psi);
    const double2 P = cmul(cmul(chi, Pd), psi);
    const double2 j0 = J[s], j1 = J[Ni + s], j2 = J[2 * Ni + s];
    const double2 jv[3] = {j0, j1, j2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        double2 g = cmul(P, cmul(d[a], d[b]));
        if (a == b) g = csub(g, Q);
        acc = cadd(acc, cmul(g, jv[b]));
      }
      e[a] = cadd(e[a], acc);
    }
  }
  E[i] = e[0];
  E[No + i] = e[1];
  E[2 * No + i] = e[2];
}

// E = Einc - s * y (Einc may be null)
__global__ void __launch_bounds__(256) k_total_field(const double2* __restrict__ einc, const double2* __restrict__ y,
                                                     double2 s, int64_t n, double2* __restrict__ E) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 v = cmul(s, y[i]);
  E[i] = einc ? csub(einc[i], v) : cneg(v);
}

void dgf_run(vp_ctx ctx, const void* J, const void* Icoord, int64_t Ni, ObsPoints obs, const double* Ocoord_dev,
             int64_t No, double ko, double cutoff, void* E, int mem) {
  VP_REQUIRE(ctx && J && Icoord && E, ERR_INVALID, "ctx/J/Icoord/E is NULL");
  VP_REQUIRE(Ni > 0 && Ni < (1 << 20) && No > 0, ERR_INVALID, "source and observation counts must be positive");
  VP_REQUIRE(ko > 0 && std::isfinite(ko) && cutoff > 0, ERR_INVALID, "ko and cutoff must be positive");
  VP_REQUIRE(mem == VP_MEM_HOST || mem == VP_MEM_DEVICE, ERR_INVALID, "mem must be VP_MEM_HOST or VP_MEM_DEVICE");
  VP_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  DevBuf dj(size_t(3 * Ni) * sizeof(double2)), dic(size_t(3 * Ni) * sizeof(double2));
  VP_HIP(hipMemcpyAsync(dj.p, J, dj.bytes, hipMemcpyHostToDevice, s));
  VP_HIP(hipMemcpyAsync(dic.p, Icoord, dic.bytes, hipMemcpyHostToDevice, s));
  DevBuf doc, de;
  if (obs.oc) {
    if (mem == VP_MEM_HOST) {
      doc.alloc(size_t(3 * No) * sizeof(double));
      VP_HIP(hipMemcpyAsync(doc.p, Ocoord_dev, doc.bytes, hipMemcpyHostToDevice, s));
      obs.oc = doc.as<double>();
    }
  }
  double2* dE = static_cast<double2*>(E);
  if (mem == VP_MEM_HOST) {
    de.alloc(size_t(3 * No) * sizeof(double2));
    dE = de.as<double2>();
  }
  const double omega = kC0 * ko;   // 2 pi c / lambda with lambda = 2 pi / ko
  hipLaunchKernelGGL(k_dgf, dim3(unsigned((No + 255) / 256)), dim3(256), 0, s, dj.as<double2>(), dic.as<double2>(),
                     int(Ni), obs, No, ko, omega, cutoff, dE);
  VP_HIP(hipGetLastError());
  if (mem == VP_MEM_HOST) VP_HIP(hipMemcpyAsync(E, dE, de.bytes, hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace vp

using namespace vp;

extern "C" {

int vp_field_dgf(vp_ctx ctx, const void* J, const void* Icoord, int64_t Ni, const double* Ocoord, int64_t No,
                 double ko, double cutoff, void* E, int mem) {
  VP_API_BEGIN
  VP_REQUIRE(Ocoord, ERR_INVALID, "Ocoord is NULL");
  ObsPoints obs{Ocoord, 0, 0, 0, 0, 0, 0};
  dgf_run(ctx, J, Icoord, Ni, obs, Ocoord, No, ko, cutoff, E, mem);
  VP_API_END
}

int vp_field_dgf_grid(vp_ctx ctx, const void* J, const void* Icoord, int64_t Ni, const double* origin, double dx,
                      int64_t L, int64_t M, int64_t N, double ko, double cutoff, void* E, int mem) {
  VP_API_BEGIN
  VP_REQUIRE(origin && dx > 0 && L > 0 && M > 0 && N > 0, ERR_INVALID, "bad grid");
  VP_REQUIRE(L * M < (int64_t(1) << 31), ERR_UNSUPPORTED, "grid plane too large");
  ObsPoints obs{nullptr, origin[0], origin[1], origin[2], dx, int(L), int(M)};
  dgf_run(ctx, J, Icoord, Ni, obs, nullptr, L * M * N, ko, cutoff, E, mem);
  VP_API_END
}

int vp_field_total(vp_op op, const void* J, const void* Einc, double freq, void* E, int mem) {
  VP_API_BEGIN
  VP_REQUIRE(op && J && E, ERR_INVALID, "op/J/E is NULL");
  VP_REQUIRE(op->all_voxels && op->unit_contrast, ERR_INVALID,
             "field operator must cover every voxel (idx = 1..3V) with Mc = 1 and no Mr");
  VP_REQUIRE(op->p64.present, ERR_INVALID, "field operator has no FP64 copy");
  VP_REQUIRE(freq > 0 && std::isfinite(freq), ERR_INVALID, "freq must be positive");
  VP_REQUIRE(mem == VP_MEM_HOST || mem == VP_MEM_DEVICE, ERR_INVALID, "mem must be VP_MEM_HOST or VP_MEM_DEVICE");
  VP_HIP(hipSetDevice(op->ctx->device));
  hipStream_t s = op->ctx->stream;
  const int64_t n = op->n;
  const size_t bytes = size_t(n) * sizeof(double2);
  DevBuf dj, dinc, de, dy(bytes);
  const double2* pj = static_cast<const double2*>(J);
  const double2* pinc = static_cast<const double2*>(Einc);
  double2* pe = static_cast<double2*>(E);
  if (mem == VP_MEM_HOST) {
    dj.alloc(bytes);
    VP_HIP(hipMemcpyAsync(dj.p, J, bytes, hipMemcpyHostToDevice, s));
    pj = dj.as<double2>();
    if (Einc) {
      dinc.alloc(bytes);
      VP_HIP(hipMemcpyAsync(dinc.p, Einc, bytes, hipMemcpyHostToDevice, s));
      pinc = dinc.as<double2>();
    }
    de.alloc(bytes);
    pe = de.as<double2>();
  }
  // y = Gram J - N(J); E = Einc + (N(J) - Gram J) / (Gram j w eps0) = Einc - y / (Gram j w eps0)
  op_apply_dev<double>(op, VP_NOTRANSP, pj, dy.as<double2>());
  const double omega = 2.0 * 3.14159265358979323846 * freq;
  const double2 sc = make_double2(0.0, -1.0 / (op->gram * omega * kEps0));   // 1 / (Gram j w eps0)
  hipLaunchKernelGGL(k_total_field, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, pinc, dy.as<double2>(), sc, n,
                     pe);
  VP_HIP(hipGetLastError());
  if (mem == VP_MEM_HOST) VP_HIP(hipMemcpyAsync(E, pe, bytes, hipMemcpyDeviceToHost, s));
  VP_HIP(hipStreamSynchronize(s));
  VP_API_END
}

}  // extern "C"
