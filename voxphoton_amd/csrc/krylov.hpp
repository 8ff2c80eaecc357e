// Restarted GMRES on the device (pgmres.m:70-180 / pgmres_iter.m:22-150
// semantics) with classical Gram-Schmidt, a selective second pass (DGKS
// criterion) and FP64 Givens least squares, batched over independent problems.
#pragma once
#include <cmath>
#include <limits>
#include <memory>
#include <vector>
#include "precond.hpp"

namespace vp {

// Relative residual |r| / |b|.  A non-finite norm (a NaN or Inf in b, x or
// the operator's output) gives NaN, which no tolerance accepts, so the
// problem stops unconverged (flag 1); b = 0 gives 0 (x = 0 solves it).
inline double relres_of(double rnorm, double bnorm) {
  if (!std::isfinite(rnorm) || !std::isfinite(bnorm)) return std::numeric_limits<double>::quiet_NaN();
  return bnorm > 0 ? rnorm / bnorm : 0.0;
}

// Progress lines on stderr (environment VP_PROGRESS=1): one per GMRES cycle
// and per refinement step, for long solves whose callers want to see them move.
bool progress_enabled();

// Device workspaces of gmres_device kept by a problem set between calls (the
// inner solves of one refinement run) with the hipGraphs of Arnoldi step
// chunks captured over them (krylov.hip).
struct GmresWork;
void destroy_gmres_work(GmresWork* w);
struct GmresWorkDeleter {
  void operator()(GmresWork* w) const { destroy_gmres_work(w); }
};

// One set of independent problems of identical geometry solved together.
// Vector p of a batch sits at base + p * n.
template <typename T>
struct ProblemSet {
  int nprob = 0;
  int64_t n = 0;
  hipStream_t stream = nullptr;
  // operator
  Geometry g{};
  MatvecPlan plan{};
  DevBuf probs;                   // ProbPtrs<T>[nprob]
  DevBuf works, works2;           // per-problem spectrum workspaces (may be empty: ops' own)
  const C<T>* twx = nullptr;
  const C<T>* twy = nullptr;
  const C<T>* twz = nullptr;
  // preconditioner (optional): one part, or the parts of block-Jacobi
  // composites applied one after the other
  struct PrecPart {
    PcGeom pg{};
    PcArgs<T> pa{};                 // x, y, active, strides filled at call time
    DevBuf pblocks, pwork, pwork2;  // block pointer array, workspaces
    DevBuf phblocks, phscale;       // binary16 block pointers and scales (FP32 quadrant form)
  };
  bool has_prec = false;
  std::vector<std::unique_ptr<PrecPart>> pparts;
  DevBuf active;                  // int[nprob]
  std::unique_ptr<GmresWork, GmresWorkDeleter> gw;   // gmres_device workspaces and step graphs
  int recycle_width = -1;         // the context's recycled-space width (vp_ctx_s::recycle_width)
};

template <typename T>
void ps_apply_A(ProblemSet<T>& ps, const C<T>* x, int64_t xs, C<T>* y, int64_t ys, const int* active);
// Problem set over operators (and preconditioners, may be NULL) of one
// geometry; throws ERR_INVALID when they differ.  A single problem uses its
// operator's own spectrum workspace unless private_work asks for new ones
// (callers that may run one operator on two streams at once).
template <typename T>
void make_problem_set(ProblemSet<T>& ps, vp_op_s* const* ops, vp_prec_s* const* precs, int np,
                      bool private_work = false);
template <typename T>
void ps_apply_M(ProblemSet<T>& ps, const C<T>* x, int64_t xs, C<T>* y, int64_t ys, const int* active);

// Recycled space kept across the inner FP32 solves of one refinement run
// (GCRO with the whole Krylov space of the first inner solve): C orthonormal
// and U with Aop U = C, Aop = (A/s) M^-1 (right) or M^-1 (A/s) (left).  C is
// stored right-aligned in the leading kcap columns of the Krylov buffer V
// (columns kcap-kc .. kcap-1), so Gram-Schmidt against [C, V] runs over
// contiguous columns; the Arnoldi basis starts at column kcap.  Problems
// whose own space is narrower carry zero columns.
struct Recycle {
  DevBuf V;            // [p][(kcap + m + 1) * n] precision of the inner solve
  DevBuf U;            // [p][kcap * n]
  int kcap = 0, m = 0;
  int kc = 0;          // recycled columns in use (0: none yet)
  bool build = false;  // form C and U from the last cycle of the next solve
  // Per-problem spaces (problem sets whose members bring recycled spaces of
  // their own, the request server): C and U are right-aligned in the kcap
  // columns, member p's own columns are the last kcp[p] (columns left of them
  // are zero, which changes no result), and want[p] asks the next solve to
  // build p's space from its last cycle while the others keep theirs.  Empty:
  // every problem has kc columns and builds together when kc == 0.
  std::vector<int> kcp, want;
  void alloc(int64_t n, int np, int m_, size_t elem);
};

struct GmresParams {
  int restart = 50;
  double tol = 1e-6;
  int maxit = 10;
  int side = 0;        // 0 left, 1 right
  bool x0_nonzero = false;
  int resvec_cap = 0;  // per problem
  std::vector<double> tol_per;   // per-problem tolerance (empty: tol for all)
  std::vector<int> active_in;    // per-problem participation (empty: all); inactive problems do no work
  bool final_residual = true;    // recompute the true residual after the last cycle
  Recycle* rec = nullptr;        // augmented solve (restart must equal rec->m)
};

struct GmresResult {
  std::vector<int> flag, iter_inner, iter_outer, total;
  std::vector<double> relres;
  std::vector<std::vector<double>> resvec;
};

// b, x: device vectors (nprob * n) in precision T.  relres is computed in
// FP64 from the T-precision data.
template <typename T>
void gmres_device(ProblemSet<T>& ps, const C<T>* b, C<T>* x, const GmresParams& prm, GmresResult& res);

// Per-problem state of a mixed-precision refinement run.
struct IrProblem {
  int steps = 0, inner = 0, flag = 1;
  double relres = 1.0;
  bool done = false;
  std::vector<double> hist;   // relres at the start of every round
};
bool ir_step_set(ProblemSet<double>& p64, ProblemSet<float>& p32, const double2* b, double2* x, double scale32,
                 double tol, int restart, double inner_tol, int inner_maxit, int max_outer, int side,
                 std::vector<IrProblem>& st, Recycle* rec = nullptr);

// Flexible refinement (FGMRES in FP64 whose preconditioner is the inner FP32
// solve): outer step j solves (A/s) z_j = v_j with the FP32 GMRES and the
// preconditioner, forms w = A z_j / s in FP64, orthogonalises it against the
// outer basis (two classical Gram-Schmidt passes) and minimises the FP64
// residual over all corrections of the cycle, x += Z y.  Cycles of outer_m
// steps restart from the true residual; every problem stops at its true
// FP64 relative residual <= tol or after max_outer inner solves.
// st[p].hist holds the residual after every inner solve (the FGMRES
// estimate inside a cycle, the true residual at its end).
void fgmres_ir_set(ProblemSet<double>& p64, ProblemSet<float>& p32, const double2* b, double2* x, double scale32,
                   double tol, int restart, double inner_tol, int inner_maxit, int max_outer, int outer_m, int side,
                   std::vector<IrProblem>& st, Recycle* rec = nullptr);

// Device vector helpers (FP64 accumulation).
template <typename T>
void dev_norms(const C<T>* x, int64_t n, int nprob, int64_t vstride, double* d_out, hipStream_t s);
void dev_residual_d(const double2* b, const double2* Ax, double2* r, int64_t n, int nprob, hipStream_t s);
void dev_scale_to_f(const double2* r, const double* d_norm, float2* out, int64_t n, int nprob, double extra,
                    hipStream_t s);
void dev_update_from_f(double2* x, const float2* d, const double* d_norm, int64_t n, int nprob, double extra,
                       const int* active, hipStream_t s);

}  // namespace vp
