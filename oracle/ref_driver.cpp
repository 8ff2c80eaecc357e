// C entry points over the reference's own C++ cubature cores (test
// infrastructure): the 4D face-pair cubature of cubature_nop_const.cpp and
// the DEMCEM singular triangle-pair integrals of direct_ws_{st,ea,va}_const.cpp
// (self term, edge adjacent, vertex adjacent).  Those sources are compiled
// from where they lie in the reference by oracle/build_ref.sh; this file only
// declares their signatures and forwards plain arrays.
#include <complex>

using cplx = std::complex<double>;

// signatures of the reference functions (cubature_nop_const.cpp:8,
// direct_ws_st_const.cpp:53, direct_ws_ea_const.cpp:58, direct_ws_va_const.cpp:62)
void cubature_nop(const int Np, const double w[], const double u[], const double v[], const double r_m[3],
                  const double r_n[3], const double R_faces[3][4][6], const double ko, const double dx,
                  cplx I_mn[6][6]);
cplx direct_ws_st_const(const double r1[], const double r2[], const double r3[], const double ko, const int Np_1D,
                        const double z[], const double w[]);
cplx direct_ws_ea_const(const double r1[], const double r2[], const double r3[], const double r4[], const double ko,
                        const int N_theta, const int N_psi, const double w_theta[], const double z_theta[],
                        const double w_psi[], const double z_psi[]);
cplx direct_ws_va_const(const double r1[], const double r2[], const double r3[], const double r4[],
                        const double r5[], const double ko, const int N_theta_p, const int N_theta_q,
                        const int N_psi, const double w_theta_p[], const double z_theta_p[],
                        const double w_theta_q[], const double z_theta_q[], const double w_psi[],
                        const double z_psi[]);

extern "C" {

// faces: R_faces[j][corner][face] as 72 doubles in that (C) order; out: 6 x 6
// complex, row-major (observation face, source face), interleaved re/im.
void vpref_cubature_nop(int Np, const double* w, const double* u, const double* v, const double* r_m,
                        const double* r_n, const double* faces, double ko, double dx, double* out) {
  double R[3][4][6];
  for (int j = 0; j < 3; ++j)
    for (int c = 0; c < 4; ++c)
      for (int f = 0; f < 6; ++f) R[j][c][f] = faces[(j * 4 + c) * 6 + f];
  cplx I[6][6];
  cubature_nop(Np, w, u, v, r_m, r_n, R, ko, dx, I);
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      out[2 * (a * 6 + b)] = I[a][b].real();
      out[2 * (a * 6 + b) + 1] = I[a][b].imag();
    }
}

void vpref_direct_ws_st(const double* r1, const double* r2, const double* r3, double ko, int np, const double* z,
                        const double* w, double* out) {
  const cplx v = direct_ws_st_const(r1, r2, r3, ko, np, z, w);
  out[0] = v.real();
  out[1] = v.imag();
}

void vpref_direct_ws_ea(const double* r1, const double* r2, const double* r3, const double* r4, double ko,
                        int n_theta, int n_psi, const double* w_theta, const double* z_theta, const double* w_psi,
                        const double* z_psi, double* out) {
  const cplx v = direct_ws_ea_const(r1, r2, r3, r4, ko, n_theta, n_psi, w_theta, z_theta, w_psi, z_psi);
  out[0] = v.real();
  out[1] = v.imag();
}

void vpref_direct_ws_va(const double* r1, const double* r2, const double* r3, const double* r4, const double* r5,
                        double ko, int n_theta_p, int n_theta_q, int n_psi, const double* w_theta_p,
                        const double* z_theta_p, const double* w_theta_q, const double* z_theta_q,
                        const double* w_psi, const double* z_psi, double* out) {
  const cplx v = direct_ws_va_const(r1, r2, r3, r4, r5, ko, n_theta_p, n_theta_q, n_psi, w_theta_p, z_theta_p,
                                    w_theta_q, z_theta_q, w_psi, z_psi);
  out[0] = v.real();
  out[1] = v.imag();
}

}  // extern "C"
