// Eigenpairs of a small dense complex upper-Hessenberg matrix on the host
// (the harmonic Ritz problem of a GMRES cycle, order <= a few hundred).
//
// Method: the complex single-shift QR algorithm on the Hessenberg form
// (Golub & Van Loan, Matrix Computations, 4th ed., algorithms 7.5.1-7.5.2,
// with the Wilkinson shift and deflation of negligible subdiagonal entries
// |h(i,i-1)| <= eps (|h(i-1,i-1)| + |h(i,i)|)), accumulating the unitary
// factor Z of the Schur form A = Z T Z^H; eigenvectors of T by back
// substitution, (T - t_kk I) y = 0 with y_k = 1 and y_j = 0 for j > k, then
// x = Z y normalised to unit 2-norm.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace vp {

using cplx = std::complex<double>;

// A: n x n column-major upper Hessenberg (entries below the first
// subdiagonal are ignored).  On success returns true with the eigenvalues in
// w (n) and unit eigenvectors in X (n x n column-major, column j for w[j]);
// false when the iteration did not converge (30 n sweeps per eigenvalue).
inline bool hessenberg_eig(int n, const std::vector<cplx>& A, std::vector<cplx>& w, std::vector<cplx>& X) {
  std::vector<cplx> H(A);
  auto h = [&](int i, int j) -> cplx& { return H[size_t(j) * n + i]; };
  for (int j = 0; j < n; ++j)
    for (int i = j + 2; i < n; ++i) h(i, j) = 0.0;
  std::vector<cplx> Z(size_t(n) * n, 0.0);
  auto z = [&](int i, int j) -> cplx& { return Z[size_t(j) * n + i]; };
  for (int i = 0; i < n; ++i) z(i, i) = 1.0;
  const double eps = 2.220446049250313e-16;
  int hi = n - 1;
  int iter = 0;
  while (hi > 0) {
    // find the active block [lo, hi]: the last negligible subdiagonal
    int lo = hi;
    while (lo > 0) {
      const double s = std::abs(h(lo - 1, lo - 1)) + std::abs(h(lo, lo));
      if (std::abs(h(lo, lo - 1)) <= eps * (s > 0 ? s : 1.0)) {
        h(lo, lo - 1) = 0.0;
        break;
      }
      --lo;
    }
    if (lo == hi) {   // 1 x 1 block converged
      --hi;
      iter = 0;
      continue;
    }
    if (++iter > 30 * n) return false;
    // Wilkinson shift: the eigenvalue of the trailing 2 x 2 closer to h(hi,hi);
    // an exceptional shift every tenth sweep breaks cycles
    const cplx a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
    cplx mu;
    if (iter % 10 == 0) {
      mu = d + cplx(std::abs(c), 0.0);
    } else {
      const cplx tr = a + d, det = a * d - b * c;
      const cplx disc = std::sqrt(tr * tr / 4.0 - det);
      const cplx l1 = tr / 2.0 + disc, l2 = tr / 2.0 - disc;
      mu = std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
    }
    // one QR sweep of H - mu I on rows/columns lo..hi by Givens rotations
    // G_k = [c s; -conj(s) c] (c real) zeroing h(k+1,k) in turn
    std::vector<double> gc(size_t(hi - lo));
    std::vector<cplx> gs(size_t(hi - lo));
    for (int i = lo; i <= hi; ++i) h(i, i) -= mu;
    for (int k = lo; k < hi; ++k) {
      const cplx x = h(k, k), y = h(k + 1, k);
      const double r = std::sqrt(std::norm(x) + std::norm(y));
      double cc = 1.0;
      cplx ss = 0.0;
      if (r > 0) {
        const double ax = std::abs(x);
        if (ax > 0) {
          cc = ax / r;
          ss = (x / ax) * std::conj(y) / r;
        } else {
          cc = 0.0;
          ss = std::conj(y) / std::abs(y);
        }
      }
      gc[size_t(k - lo)] = cc;
      gs[size_t(k - lo)] = ss;
      // rows k, k+1 from column k on (columns of the whole matrix so that the
      // Schur form T stays consistent with Z)
      for (int j = k; j < n; ++j) {
        const cplx u = h(k, j), v = h(k + 1, j);
        h(k, j) = cc * u + ss * v;
        h(k + 1, j) = -std::conj(ss) * u + cc * v;
      }
    }
    // R Q: columns k, k+1 times G_k^H, rows 0 .. min(k+2, hi)
    for (int k = lo; k < hi; ++k) {
      const double cc = gc[size_t(k - lo)];
      const cplx ss = gs[size_t(k - lo)];
      const int rmax = std::min(k + 2, hi);
      for (int i = 0; i <= rmax; ++i) {
        const cplx u = h(i, k), v = h(i, k + 1);
        h(i, k) = cc * u + std::conj(ss) * v;
        h(i, k + 1) = -ss * u + cc * v;
      }
      for (int i = 0; i < n; ++i) {
        const cplx u = z(i, k), v = z(i, k + 1);
        z(i, k) = cc * u + std::conj(ss) * v;
        z(i, k + 1) = -ss * u + cc * v;
      }
    }
    for (int i = lo; i <= hi; ++i) h(i, i) += mu;
  }
  // eigenvalues: diagonal of the Schur form; eigenvectors by back substitution
  w.assign(size_t(n), 0.0);
  for (int i = 0; i < n; ++i) w[size_t(i)] = h(i, i);
  double tnorm = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) tnorm = std::max(tnorm, std::abs(h(i, j)));
  const double small = std::max(eps * tnorm, 1e-300);
  X.assign(size_t(n) * n, 0.0);
  std::vector<cplx> y(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) {
    std::fill(y.begin(), y.end(), cplx(0.0));
    y[size_t(k)] = 1.0;
    const cplx lam = h(k, k);
    for (int i = k - 1; i >= 0; --i) {
      cplx s = 0.0;
      for (int j = i + 1; j <= k; ++j) s += h(i, j) * y[size_t(j)];
      cplx dd = h(i, i) - lam;
      if (std::abs(dd) < small) dd = small;   // equal eigenvalues: perturbed pivot
      y[size_t(i)] = -s / dd;
    }
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      cplx s = 0.0;
      for (int j = 0; j <= k; ++j) s += z(i, j) * y[size_t(j)];
      X[size_t(k) * n + i] = s;
      nrm += std::norm(s);
    }
    nrm = std::sqrt(nrm);
    if (nrm > 0)
      for (int i = 0; i < n; ++i) X[size_t(k) * n + i] /= nrm;
  }
  return true;
}

// Recycled pair of one problem from its last cycle's Arnoldi relation
// Aop V_K = V_{K+1} Hb (Hb (K+1) x K, column-major): the kd harmonic Ritz
// vectors of smallest modulus, as GMRES-DR selects them (Morgan, SIAM J.
// Sci. Comput. 24 (2002); GCRO-DR, Parks et al., SIAM J. Sci. Comput. 28
// (2006)): eigenvectors P of H_K + |h_{K+1,K}|^2 H_K^-H e_K e_K^H, then
// Hb P = Q2 R2 (QR) gives C = V_{K+1} Q2 orthonormal and U = V_K P R2^-1
// with Aop U = C.  sc ((K+1) x kd) and su (K x kd), column-major, get the
// coefficients; returns the columns kept (0: none, the caller keeps the
// whole cycle).
inline int harmonic_pair(int K, const std::vector<cplx>& Hb, int kd, std::vector<cplx>& sc, std::vector<cplx>& su) {
  auto hb = [&](int i, int j) { return Hb[size_t(j) * (K + 1) + i]; };
  // f = H_K^-H e_K by Gaussian elimination with partial pivoting
  std::vector<cplx> Mt(size_t(K) * K), f(size_t(K), 0.0);
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) Mt[size_t(j) * K + i] = std::conj(hb(j, i));   // (H_K^H)(i, j)
  f[size_t(K - 1)] = 1.0;
  for (int c = 0; c < K; ++c) {
    int piv = c;
    for (int r = c + 1; r < K; ++r)
      if (std::abs(Mt[size_t(c) * K + r]) > std::abs(Mt[size_t(c) * K + piv])) piv = r;
    if (std::abs(Mt[size_t(c) * K + piv]) == 0.0) return 0;
    if (piv != c) {
      for (int j = 0; j < K; ++j) std::swap(Mt[size_t(j) * K + c], Mt[size_t(j) * K + piv]);
      std::swap(f[size_t(c)], f[size_t(piv)]);
    }
    const cplx d = Mt[size_t(c) * K + c];
    for (int r = c + 1; r < K; ++r) {
      const cplx l = Mt[size_t(c) * K + r] / d;
      if (l == cplx(0.0)) continue;
      for (int j = c; j < K; ++j) Mt[size_t(j) * K + r] -= l * Mt[size_t(j) * K + c];
      f[size_t(r)] -= l * f[size_t(c)];
    }
  }
  for (int r = K - 1; r >= 0; --r) {
    cplx t = f[size_t(r)];
    for (int j = r + 1; j < K; ++j) t -= Mt[size_t(j) * K + r] * f[size_t(j)];
    f[size_t(r)] = t / Mt[size_t(r) * K + r];
  }
  const double h2 = std::norm(hb(K, K - 1));
  std::vector<cplx> G(size_t(K) * K);
  for (int j = 0; j < K; ++j)
    for (int i = 0; i < K; ++i) G[size_t(j) * K + i] = hb(i, j);
  for (int i = 0; i < K; ++i) G[size_t(K - 1) * K + i] += h2 * f[size_t(i)];
  std::vector<cplx> w, X;
  if (!hessenberg_eig(K, G, w, X)) return 0;
  std::vector<int> ord(static_cast<size_t>(K));
  for (int i = 0; i < K; ++i) ord[size_t(i)] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return std::abs(w[size_t(a)]) < std::abs(w[size_t(b)]); });
  kd = std::min(kd, K);
  // Hb P, then QR by modified Gram-Schmidt with a second pass
  std::vector<cplx> Y(size_t(K + 1) * kd, 0.0), Pm(size_t(K) * kd);
  for (int c = 0; c < kd; ++c) {
    const cplx* x = X.data() + size_t(ord[size_t(c)]) * K;
    for (int i = 0; i < K; ++i) Pm[size_t(c) * K + i] = x[i];
    for (int j = 0; j < K; ++j)
      for (int i = 0; i <= std::min(j + 1, K); ++i) Y[size_t(c) * (K + 1) + i] += hb(i, j) * x[j];
  }
  std::vector<cplx> R2(size_t(kd) * kd, 0.0);
  int kept = 0;
  for (int c = 0; c < kd; ++c) {
    cplx* y = Y.data() + size_t(c) * (K + 1);
    double n0 = 0.0;
    for (int i = 0; i <= K; ++i) n0 += std::norm(y[i]);
    n0 = std::sqrt(n0);
    for (int pass = 0; pass < 2; ++pass)
      for (int q = 0; q < kept; ++q) {
        const cplx* u = Y.data() + size_t(q) * (K + 1);
        cplx d = 0.0;
        for (int i = 0; i <= K; ++i) d += std::conj(u[i]) * y[i];
        for (int i = 0; i <= K; ++i) y[i] -= d * u[i];
        R2[size_t(c) * kd + q] += d;
      }
    double nr = 0.0;
    for (int i = 0; i <= K; ++i) nr += std::norm(y[i]);
    nr = std::sqrt(nr);
    if (!(nr > 1e-12 * n0) || !(nr > 0)) break;   // dependent direction: keep the columns so far
    for (int i = 0; i <= K; ++i) y[i] /= nr;
    R2[size_t(c) * kd + c] = nr;
    if (c != kept) return 0;
    ++kept;
  }
  if (kept == 0) return 0;
  sc.assign(size_t(K + 1) * kept, 0.0);
  su.assign(size_t(K) * kept, 0.0);
  for (int c = 0; c < kept; ++c)
    for (int i = 0; i <= K; ++i) sc[size_t(c) * (K + 1) + i] = Y[size_t(c) * (K + 1) + i];
  // su = P R2^-1, column by column
  for (int c = 0; c < kept; ++c) {
    for (int i = 0; i < K; ++i) {
      cplx t = Pm[size_t(c) * K + i];
      for (int q = 0; q < c; ++q) t -= su[size_t(q) * K + i] * R2[size_t(c) * kd + q];
      su[size_t(c) * K + i] = t / R2[size_t(c) * kd + c];
    }
  }
  return kept;
}

}  // namespace vp
