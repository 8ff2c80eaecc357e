// Test shim: exposes vp::hessenberg_eig (csrc/eig.hpp) to the CPU tests.
#include "eig.hpp"

extern "C" int eig_shim(int n, const double* a, double* w, double* x) {
  std::vector<vp::cplx> A(size_t(n) * n), W, X;
  for (size_t i = 0; i < A.size(); ++i) A[i] = vp::cplx(a[2 * i], a[2 * i + 1]);
  if (!vp::hessenberg_eig(n, A, W, X)) return 1;
  for (int i = 0; i < n; ++i) { w[2 * i] = W[i].real(); w[2 * i + 1] = W[i].imag(); }
  for (size_t i = 0; i < X.size(); ++i) { x[2 * i] = X[i].real(); x[2 * i + 1] = X[i].imag(); }
  return 0;
}

// harmonic_pair: hb (K+1) x K, sc (K+1) x kd, su K x kd, column-major interleaved complex
extern "C" int harm_shim(int K, const double* hb, int kd, double* sc, double* su) {
  std::vector<vp::cplx> Hb(size_t(K + 1) * K), SC, SU;
  for (size_t i = 0; i < Hb.size(); ++i) Hb[i] = vp::cplx(hb[2 * i], hb[2 * i + 1]);
  const int kept = vp::harmonic_pair(K, Hb, kd, SC, SU);
  for (size_t i = 0; i < SC.size(); ++i) { sc[2 * i] = SC[i].real(); sc[2 * i + 1] = SC[i].imag(); }
  for (size_t i = 0; i < SU.size(); ++i) { su[2 * i] = SU[i].real(); su[2 * i + 1] = SU[i].imag(); }
  return kept;
}
