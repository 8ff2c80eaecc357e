/*
 * voxphoton.h - C ABI of the MI355X-native FFT-VIE solve path.
 *
 * Replaces, for one GPU, the three function handles VoxPhoton's MATLAB
 * scripts hand to GMRES and the GMRES call itself:
 *   MVP_A  = @(J) mv_AN_const(J,fN,Mr,Mc,Gram,'notransp',idxS3,GPU_flag)
 *   prec_A = @(J) chan_mvp_idx(...) / chan_2_mvp_for_parallel_idx(...)
 *   [x,flag,relres,iter,resvec] = gmres(afun,b,restart,tol,maxit,mfun)
 * (Bragg_MOR_script.m:78-90, resonator_MOR_scripts.m:197-208,
 *  JVIE_Solver_const.m:282,301-333).
 *
 * Conventions
 *  - Complex numbers are interleaved (re, im) pairs: double[2*n] for
 *    complex128, float[2*n] for complex64 (MATLAB R2018a interleaved layout).
 *  - Grid arrays are column-major (MATLAB order): voxel (l,m,n), 0-based,
 *    sits at l + L*(m + M*n); component c of a 3-vector field at c*V + voxel.
 *  - Index vectors are 1-based int64 (the MEX gateway converts MATLAB's
 *    double indices); idx addresses the 3*L*M*N component-major array.
 *  - Every entry point returns a status (VP_OK = 0, negative on error);
 *    vp_last_error() returns a thread-local message for the last failure.
 *    No C++ exception crosses the ABI.
 *  - Handles are opaque and freed explicitly.  Calls on one handle are
 *    serialised on its context's HIP stream; separate contexts are
 *    independent.  mem = VP_MEM_HOST means the pointers are host memory
 *    (always complex128 for vectors); VP_MEM_DEVICE means device memory in
 *    the precision named by prec.
 */
#ifndef VOXPHOTON_H
#define VOXPHOTON_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VP_OK 0
#define VP_ERR_INVALID (-1)
#define VP_ERR_OOM (-2)
#define VP_ERR_DEVICE (-3)
#define VP_ERR_UNSUPPORTED (-4)

#define VP_MEM_HOST 0
#define VP_MEM_DEVICE 1

#define VP_FP64 1
#define VP_FP32 2

#define VP_NOTRANSP 0
#define VP_TRANSP 1

#define VP_PREC_LEFT 0
#define VP_PREC_RIGHT 1

typedef struct vp_ctx_s* vp_ctx;
typedef struct vp_op_s* vp_op;
typedef struct vp_prec_s* vp_prec;

/* ---- library / context --------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char* vp_last_error(void);
/* Library version string. */
const char* vp_version(void);
/* Number of visible GPUs (0 if none). */
int vp_device_count(int* count);
/* Context: one device and one HIP stream (default priority). */
int vp_ctx_create(int device, vp_ctx* out);
int vp_ctx_destroy(vp_ctx ctx);
int vp_ctx_synchronize(vp_ctx ctx);
/* Per-kernel timing on the context's stream with HIP events: after
 * vp_ctx_set_profiling(ctx, 1) every launch of the named kernels ("xfwd",
 * "plane" (fused yz pass) or "yfwd"/"zpass"/"yinv" (split yz passes), "xinv",
 * "pc_xfwd", "pc_plane" (fused preconditioner plane) or "pc_yrows"/"pc_block"
 * (row transforms and block products of the split and quadrant forms),
 * "pc_xinv") records an event pair;
 * vp_ctx_kernel_stats returns the launch count and summed milliseconds
 * (reset != 0 clears them).  No counterpart in the reference (which only
 * prints tic/toc totals); used by the benchmark's roofline figures. */
int vp_ctx_set_profiling(vp_ctx ctx, int on);
/*
 * Recycled Krylov space of the mixed-precision solves on this context
 * (vp_solve_ir*, vp_solve_fgmres_ir*, the request server): the first inner
 * solve's last cycle becomes the space the later inner solves are augmented
 * with.  width 0 keeps the whole cycle; width k > 0 keeps the k harmonic Ritz
 * directions of smallest modulus when the cycle is longer (GCRO-DR's
 * selection); -1 (the default) takes VP_IR_RECYCLE_K from the environment,
 * else 0.  The request server's rounds use the setting of the context the
 * server was created on.  A long cycle kept whole makes every later Arnoldi step
 * orthogonalise against all of it; a narrow space saves that work when few
 * inner solves follow.  No counterpart in the reference.
 */
int vp_ctx_set_recycle_width(vp_ctx ctx, int width);
int vp_ctx_kernel_stats(vp_ctx ctx, const char* kernel, int64_t* launches, double* total_ms, int reset);
/* Free and total device memory of the context's GPU (bytes); the free figure
 * counts the library's cache of released buffers as used. */
int vp_mem_info(vp_ctx ctx, int64_t* free_bytes, int64_t* total_bytes);
/* Device memory helpers for callers without their own allocator. */
int vp_malloc(vp_ctx ctx, int64_t bytes, void** dptr);
int vp_free(vp_ctx ctx, void* dptr);
int vp_memcpy_h2d(vp_ctx ctx, void* dst, const void* src, int64_t bytes);
int vp_memcpy_d2h(vp_ctx ctx, void* dst, const void* src, int64_t bytes);

/* ---- N operator (replaces mv_AN_const, pc/src_vie/src_mvp/mv_AN_const.m:1) -- */

/*
 * Creates an operator from the reference's precomputed circulant Green
 * tensor fN (getOPERATORS_SAM.m:1 first output; Lp x Mp x Np x 6 complex128,
 * column-major, components xx,xy,xz,yy,yz,zz).  Mr, Mc: L x M x N complex128
 * (Mr may be NULL for Mr == 1, the JVIE-II case).  gram: the real Gram factor
 * dx^3.  idx: n one-based indices into the 3*L*M*N field (idxS3).
 * precisions: VP_FP64, VP_FP32 or both OR-ed; FP32 copies are stored
 * prescaled by 1/gram (see vp_op_fp32_scale).  The matvec runs at the
 * tensor's own sizes Lp x Mp x Np, which must be at least 2B-1 for the
 * bounding box B of the unknowns' voxels (the reference's 2L x 2M x 2N are).
 */
int vp_op_create(vp_ctx ctx, int64_t L, int64_t M, int64_t N, int64_t Lp, int64_t Mp, int64_t Np,
                 const double* fN, const double* Mr, const double* Mc, double gram, const int64_t* idx,
                 int64_t n, int precisions, vp_op* out);

/*
 * Same operator from the Toeplitz tensor opToeplitz (getOPERATORS_SAM.m:1
 * second output; L x M x N x 6 complex128).  The circulant embedding and
 * its FFT (circulant_nop_const.m, fft_operator.m) run on the device, on the
 * offsets inside the bounding box B of the unknowns' voxels: any circulant
 * of at least 2B-1 points per axis gives the same operator on idx as the
 * reference's 2L x 2M x 2N.  A size of 0 lets an autotuner choose that
 * axis per precision (timed on the device, cached per box; environment
 * VP_FFT_SIZES=reference selects 2L x 2M x 2N); given sizes must be at least
 * 2B-1.  vp_op_info reports the sizes chosen.
 */
int vp_op_create_toeplitz(vp_ctx ctx, int64_t L, int64_t M, int64_t N, int64_t Lp, int64_t Mp, int64_t Np,
                          const double* G, const double* Mr, const double* Mc, double gram, const int64_t* idx,
                          int64_t n, int precisions, vp_op* out);

/*
 * Assembles the Toeplitz tensor on the device for a uniform grid of voxel
 * size dx at wavenumber ko (assembly_nop_const.m:1-168 semantics, own
 * cubature; only the offsets inside the unknowns' bounding box) and creates
 * the operator from it, FFT sizes as vp_op_create_toeplitz.  gram is dx^3.
 */
int vp_op_create_assembled(vp_ctx ctx, int64_t L, int64_t M, int64_t N, int64_t Lp, int64_t Mp, int64_t Np,
                           double dx, double ko, const double* Mr, const double* Mc, const int64_t* idx,
                           int64_t n, int precisions, vp_op* out);

/* Writes the assembled L x M x N x 6 Toeplitz tensor (complex128) to G (mem). */
int vp_assemble_toeplitz(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko, double* G, int mem);

/*
 * y = A x ('notransp') or A^H x ('transp'), A = Gram*Mr - Mc*N restricted to
 * idx (mv_AN_const.m:66-146).  mem = VP_MEM_HOST: x, y are complex128 host
 * vectors of length n and the result is A x for either prec.  mem =
 * VP_MEM_DEVICE: x, y are device vectors in precision prec; for VP_FP32 the
 * result is A x / vp_op_fp32_scale.  x and y must not alias.
 */
int vp_op_apply(vp_op op, int transp, const void* x, void* y, int mem, int prec);
/* Batched apply: nvec device vectors (stride n elements) through ops[i]
 * (same geometry), one launch per pass. */
int vp_op_apply_batched(const vp_op* ops, int nvec, int transp, const void* x, void* y, int prec);
int vp_op_fp32_scale(vp_op op, double* scale);
int vp_op_size(vp_op op, int64_t* n);
/*
 * Geometry and plans of an operator, 17 entries: out[0..2] the grid L, M, N;
 * out[3..5] the origin and out[6..8] the extents of the unknowns' bounding
 * box; out[9..11] the FFT sizes of the FP64 plan and out[12..14] of the FP32
 * plan (0 when that precision is absent); out[15], out[16] their matvec
 * variants (0 fused planes, 1 split passes, -1 absent).  No counterpart in
 * the reference, whose FFT sizes are always 2L x 2M x 2N
 * (circulant_nop_const.m:5).
 */
int vp_op_info(vp_op op, int64_t* out);
int vp_op_destroy(vp_op op);

/* ---- preconditioners (pc/src_preconditioning) -------------------------------- */

/*
 * 2-level block-circulant preconditioner apply (chan_2_mvp_for_parallel_idx.m:2-23)
 * from precomputed inverse blocks: blocks holds L*M complex128 matrices of
 * size 3N x 3N, column-major, cell {i,j} (1-based) at offset
 * ((i-1) + L*(j-1)) * 9*N*N.  idx indexes the 3*L*M*N box field.
 */
int vp_prec_create_chan2(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* blocks, const int64_t* idx,
                         int64_t n, int precisions, vp_prec* out);
/* 1-level apply (chan_mvp_idx.m:2-16): L matrices of size 3MN x 3MN. */
int vp_prec_create_chan1(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* blocks, const int64_t* idx,
                         int64_t n, int precisions, vp_prec* out);
/*
 * 2-level setup on the device from opToeplitz (level_2_parallel_func.m:1-192
 * semantics, built without the 3MN x 3MN intermediate): Mc is a scalar
 * (mc_re, mc_im) contrast, Mr(1) = mr0.  Blocks are inverted on the device.
 */
int vp_prec_create_level2(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* G, double mc_re, double mc_im,
                          double mr0, double gram, const int64_t* idx, int64_t n, int precisions, vp_prec* out);
/* Same, from a device-assembled tensor (see vp_op_create_assembled).  damping >= 0
 * assembles the preconditioner's kernel with the complex-shifted wavenumber
 * ko*(1 - j*damping), which keeps every circulant block away from singularity
 * (0 reproduces level_2_parallel_func exactly). */
int vp_prec_create_level2_assembled(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko, double damping,
                                    double mc_re,
                                    double mc_im, double mr0, const int64_t* idx, int64_t n, int precisions,
                                    vp_prec* out);
/*
 * 1-level setup on the device (level_1_parallel_func.m:1-169): for every kx
 * the 3MN x 3MN block of the x-Chan circulant of -Mc.*opToeplitz (Gram*Mr(1)
 * at the origin), exact in y and z, inverted (Gauss-Jordan on the device up
 * to 256 rows, rocSOLVER's batched LU above).  Mc is the scalar (mc_re,
 * mc_im), or, when the Mc pointer is not NULL, an L x M x N contrast field
 * (interleaved complex128, column-major) scaling the tensor at each offset.
 * The blocks are applied as vp_prec_create_chan1's (chan_mvp_idx.m).  All L
 * blocks of (3MN)^2 complex128 are kept: sized for boxes with 3MN in the
 * hundreds.  The _assembled form assembles G on the device (damping as for
 * vp_prec_create_level2_assembled, gram = dx^3).
 */
int vp_prec_create_level1(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* G, double mc_re, double mc_im,
                          const double* Mc, double mr0, double gram, const int64_t* idx, int64_t n, int precisions,
                          vp_prec* out);
int vp_prec_create_level1_assembled(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko,
                                    double damping, double mc_re, double mc_im, const double* Mc, double mr0,
                                    const int64_t* idx, int64_t n, int precisions, vp_prec* out);
/* Face-averaged 1-level form of the reference's Bragg script
 * (level_1_parallel_func_N.m:1-192 and Bragg_MOR_script.m:61-72): blocks
 * inv(Gram*I - diag(face3) * circ_N{kx}) of the bare assembled tensor, with
 * face[m + M*n] (complex128, M x N) and face3 = [face; face; face]. */
int vp_prec_create_level1_face(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko, double damping,
                               const double* face, const int64_t* idx, int64_t n, int precisions, vp_prec* out);
/*
 * 2-level setup with a contrast field, as level_2_parallel_func.m:5-75 takes
 * an L x M x N array Mc (resonator_MOR_scripts.m:165-170 passes the disk
 * square's (e_r - 1)/e_r): the Toeplitz entry at offset (i, j, k) is scaled by
 * Mc(i, j, k) before the blocks are formed.  Mc is interleaved complex128 in
 * column-major order; Mr(1) = mr0.  The _assembled form assembles G on the
 * device (damping as for vp_prec_create_level2_assembled).
 */
int vp_prec_create_level2_field(vp_ctx ctx, int64_t L, int64_t M, int64_t N, const double* G, const double* Mc,
                                double mr0, double gram, const int64_t* idx, int64_t n, int precisions, vp_prec* out);
int vp_prec_create_level2_assembled_field(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko,
                                          double damping, const double* Mc, double mr0, const int64_t* idx, int64_t n,
                                          int precisions, vp_prec* out);
/*
 * Face-averaged 2-level preconditioner (the form Bragg_MOR_script.m:61-72
 * builds with level_1_parallel_func_N: inv(Gram*I - diag(face3)*circ_N)),
 * with the circulant approximation taken along the first two axes of the box
 * and the third axis exact: blocks inv(Gram*I - diag(face(n)) * N2(kx, ky)),
 * N2 the 2-level Chan circulant of the bare N operator (Green tensor
 * assembled on the device for dx, ko), face: complex[N] contrast profile
 * along the exact axis (e.g. (e-1)/e of the box-averaged permittivity).
 * Callers permute axes so that the exact axis comes last.
 */
int vp_prec_create_level2_face(vp_ctx ctx, int64_t L, int64_t M, int64_t N, double dx, double ko, double damping,
                               const double* face,
                               const int64_t* idx, int64_t n, int precisions, vp_prec* out);
/* Embeds a preconditioner built on a sub-box into the operator's unknowns:
 * sub_of_op[i] (1-based, length nsub) is the position in the operator's
 * unknown vector of the sub-box unknown i; unknowns not covered pass
 * through unchanged (block-Jacobi, prec_fun_disk_idx.m:1-23). */
int vp_prec_set_embedding(vp_prec p, int64_t n_outer, const int64_t* sub_of_op, int64_t nsub);
/*
 * Block-Jacobi composite (prec_fun_disk_idx.m:1-23): y = x on the unknowns no
 * part covers, and on each part's unknowns that part's M^{-1} applied to
 * them.  Every part carries an embedding (vp_prec_set_embedding) into the
 * same outer vector, and the parts cover disjoint unknowns (VP_ERR_INVALID
 * otherwise).  The composite refers to the parts without owning them: they
 * must outlive it.  The reference's resonator setup (resonator_MOR_scripts.m:
 * 104-109, 165-170, 195-201) pairs a preconditioner on the bus waveguide's
 * box with a 2-level one on the disk's box; it negates both inverses, which
 * changes no GMRES iterate, so the parts here keep their own sign.
 */
int vp_prec_create_composite(vp_ctx ctx, const vp_prec* parts, int nparts, vp_prec* out);
/* y = M^{-1} x; same mem/prec conventions as vp_op_apply (the FP32 copy of a
 * preconditioner built for the FP64 operator is scaled by the op's 1/gram so
 * that it inverts the FP32 operator). */
int vp_prec_apply(vp_prec p, const void* x, void* y, int mem, int prec);
int vp_prec_set_fp32_scale(vp_prec p, double scale);
int vp_prec_destroy(vp_prec p);

/* ---- GMRES (pgmres.m:70-180, pgmres_iter.m:22-150) ----------------------- */

typedef struct {
  int flag;          /* 0 converged (relres <= tol), 1 not converged */
  int iter_inner;    /* inner iterations of the last cycle (iter(1)) */
  int iter_outer;    /* restart cycles (iter(2)) */
  int total_iters;   /* total Arnoldi steps */
  double relres;     /* ||b - A x|| / ||b||, FP64, unpreconditioned */
  double* resvec;    /* caller buffer for the residual history or NULL */
  int resvec_cap;    /* capacity of resvec */
  int resvec_len;    /* entries written */
  int ir_steps;      /* mixed precision: outer refinement steps */
  double seconds;    /* wall time of the solve */
} vp_gmres_info;

/*
 * Restarted GMRES(restart) on A x = b with optional preconditioner applied
 * on the given side.  Arnoldi uses one classical Gram-Schmidt pass per
 * step, as pgmres_iter.m:106-107 (VP_DGKS_ETA=eta adds a second pass where
 * ||w'|| < eta ||w||; VP_GMRES_CGS2=1 adds it every step); the Hessenberg
 * least squares is updated with
 * Givens rotations in FP64.  prec selects the arithmetic (VP_FP64 or VP_FP32
 * for the whole Krylov loop).  b, x0 (nullable), x follow mem.
 */
int vp_gmres(vp_op op, vp_prec M, int side, const void* b, const void* x0, void* x, int mem, int prec, int restart,
             double tol, int maxit, vp_gmres_info* info);

/*
 * Preconditioned BiCGSTAB: the 'B' solver of the reference's harness
 * (JVIE_Solver_const.m:469-494, MATLAB's bicgstab(A,b,tol,maxit,M)), in the
 * published form of van der Vorst (1992) with M (nullable) applied to the
 * search directions.  Stops when ||b - A x|| <= tol ||b|| (a candidate from
 * the recurrence is confirmed on the true residual) or after maxit
 * iterations.  info: flag 0 converged, 1 not converged, 4 breakdown
 * (<r0, r> or omega zero); relres the final ||b - A x|| / ||b||; resvec the
 * relative residual at the start and after every half iteration;
 * total_iters the half iterations, iter_inner the whole iterations and
 * iter_outer 1 when it stopped at a half iteration (MATLAB's iter = 3.5).
 * b, x0 (nullable), x, mem and prec as vp_gmres.
 */
int vp_bicgstab(vp_op op, vp_prec M, const void* b, const void* x0, void* x, int mem, int prec, double tol,
                int maxit, vp_gmres_info* info);

/*
 * Mixed precision: FP64 iterative refinement around FP32 GMRES.  Outer
 * residual r = b - A x in FP64 on the device; inner FP32 GMRES(restart)
 * with the FP32 preconditioner solves A d = r to inner_tol, or to
 * need = min(0.5, 0.5 * tol / relres) when need >= inner_tol / 8 (a step
 * that can reach tol goes as deep as it needs instead of leaving a short
 * extra step); stops when ||b - A x|| / ||b|| <= tol or after max_outer
 * steps.  The inner solves after the first are augmented with the first
 * one's Krylov space (GCRO: C = (A M^-1) U orthonormal, the residual is
 * projected on C and Arnoldi runs orthogonal to C), which keeps what the
 * first solve learnt about the slow modes; VP_IR_RECYCLE=0 restarts every
 * inner solve afresh.  The recycled space holds up to restart vectors (plus
 * U) per solve in FP32.  No counterpart in the reference (which solves in
 * FP64 only).
 */
int vp_solve_ir(vp_op op, vp_prec M, int side, const void* b, void* x, int mem, double tol, int restart,
                double inner_tol, int inner_maxit, int max_outer, vp_gmres_info* info);

/*
 * Batched mixed-precision solves: nsolve right-hand sides, solve i through
 * ops[i] and precs[i] (precs may be NULL or hold NULLs; all ops share one
 * geometry).  b and x are device complex128 arrays of nsolve*n entries.
 * Every Krylov step of all active solves runs as one launch per kernel.
 */
int vp_solve_ir_batched(const vp_op* ops, const vp_prec* precs, int nsolve, int side, const void* b, void* x,
                        double tol, int restart, double inner_tol, int inner_maxit, int max_outer,
                        vp_gmres_info* infos);

/*
 * Flexible refinement: vp_solve_ir's inner FP32 solves (same restart,
 * inner_tol, need rule and recycled Krylov space) used as the variable
 * preconditioner of an FP64 flexible GMRES (FGMRES).  Outer step j solves
 * (A/s) z_j = v_j in FP32, forms w = A z_j / s in FP64, orthogonalises it
 * against the outer basis with two classical Gram-Schmidt passes and
 * minimises ||b - A x|| over every correction of the cycle (x += Z y), where
 * plain refinement takes each correction with weight one.  Cycles of
 * outer_restart steps restart from the true FP64 residual; a solve stops at
 * ||b - A x|| / ||b|| <= tol (recomputed in FP64) or after max_outer inner
 * solves.  The history (info->resvec) holds the relative residual after
 * every inner solve: the FGMRES estimate inside a cycle, the recomputed
 * residual at its end.  Costs one FP64 matvec more per cycle than
 * vp_solve_ir, and (outer_restart + 1) complex128 plus outer_restart
 * complex64 vectors per solve.  No counterpart in the reference (which
 * solves in FP64 only).
 */
int vp_solve_fgmres_ir(vp_op op, vp_prec M, int side, const void* b, void* x, int mem, double tol, int restart,
                       double inner_tol, int inner_maxit, int max_outer, int outer_restart, vp_gmres_info* info);

/* Batched flexible refinement (vp_solve_ir_batched's arguments and layout). */
int vp_solve_fgmres_ir_batched(const vp_op* ops, const vp_prec* precs, int nsolve, int side, const void* b,
                               void* x, double tol, int restart, double inner_tol, int inner_maxit, int max_outer,
                               int outer_restart, vp_gmres_info* infos);

/* ---- request server (continuous batching) ----------------------------------
 * No counterpart in the reference, whose scripts solve one wavelength after
 * another (Bragg_MOR_script.m:88-95, example_Bragg_1500nm_to_1600nm.m:85-114).
 * A server owns one HIP stream on the context's device and a worker thread.
 * Submitted mixed-precision solves (vp_solve_ir semantics: each request
 * keeps the recycled Krylov space of its first inner solve on the device
 * between rounds, VP_SERVER_RECYCLE=0 restarts every inner solve afresh) wait
 * in a FIFO queue and are admitted while the server's HBM budget (solver state
 * of the admitted requests: Krylov basis, recycled space, refinement vectors,
 * batch workspaces) and
 * its slot count allow.  Every round, the worker groups the admitted
 * requests by operator geometry and solve parameters and runs one
 * refinement step for each group as one batched launch sequence; requests
 * admitted while others are running join their group at the next round, and
 * finished requests free their memory at once.  A request's arithmetic does
 * not depend on its group: results are identical alone and in the mix.
 * b and x are device complex128 vectors of the operator's length; the caller
 * keeps b, x, the operator and the preconditioner alive and does not use
 * them until the request is done.  Rounds keep the recycled space as the
 * creating context's vp_ctx_set_recycle_width says; that context outlives
 * the server.
 */
typedef struct vp_server_s* vp_server;
typedef struct {
  double tol;         /* outer FP64 relative residual target */
  int restart;        /* inner GMRES restart length */
  double inner_tol;   /* inner relative tolerance */
  int inner_maxit;    /* inner restart cycles per refinement step */
  int max_outer;      /* refinement steps */
  int side;           /* VP_PREC_LEFT or VP_PREC_RIGHT */
} vp_solve_params;
int vp_server_create(vp_ctx ctx, int64_t hbm_budget_bytes, int max_active, vp_server* out);
/* Queues a solve; returns at once with a ticket (> 0). */
int vp_server_submit(vp_server s, vp_op op, vp_prec M, const void* b, void* x, const vp_solve_params* prm,
                     int64_t* ticket);
/* done = 1 when the request has finished; info (nullable) then holds flag,
 * relres, ir_steps, total_iters and seconds = submit-to-finish latency
 * (resvec: relative residual at the start of every round). */
int vp_server_poll(vp_server s, int64_t ticket, int* done, vp_gmres_info* info);
/* Blocks until the request has finished; releases the ticket. */
int vp_server_wait(vp_server s, int64_t ticket, vp_gmres_info* info);
/* Queue length, admitted requests, admitted bytes and rounds run so far. */
int vp_server_stats(vp_server s, int64_t* pending, int64_t* active, int64_t* bytes_in_use, int64_t* rounds);
/* Finishes every queued request, then stops the worker. */
int vp_server_destroy(vp_server s);

/*
 * Excitation and post-processing either side of the solve (FP64).
 *
 * vp_field_dgf replaces E_field_DGF_mollified.m:1-111: the field of Ni point
 * currents J (3*Ni complex128, column-major Ni x 3, host) at source points
 * Icoord (Ni x 3 complex128, column-major, host; complex coordinates give
 * the complex-point-source beam of example_Bragg_1500nm_to_1600nm.m:64-67)
 * observed at No points Ocoord (No x 3 double, column-major, follows mem):
 * E (No x 3 complex128, column-major, follows mem) = dyadic Green's function
 * times J, weighted by the mollifier exp(-cutoff / (cutoff - R^2)) where
 * Re(R) < cutoff and 0 elsewhere (R the principal root of R.R).  The
 * reference weighs the summed field and is defined for Ni = 1; here each
 * source's term is weighed, which is the same for Ni = 1.
 * vp_field_dgf_grid observes at voxel centres origin + (l, m, n) dx of an
 * L x M x N grid (E: L*M*N x 3 in column-major voxel order), the Ocoord of
 * the configs, without materialising the coordinates.
 */
int vp_field_dgf(vp_ctx ctx, const void* J, const void* Icoord, int64_t Ni, const double* Ocoord, int64_t No,
                 double ko, double cutoff, void* E, int mem);
int vp_field_dgf_grid(vp_ctx ctx, const void* J, const void* Icoord, int64_t Ni, const double* origin, double dx,
                      int64_t L, int64_t M, int64_t N, double ko, double cutoff, void* E, int mem);
/*
 * Total field (E_field_Nop_const.m:76-84 over mv_L_const.m:61-85):
 * E = Einc + (N(J) - Gram J) / (Gram j w eps0) at every voxel, w = 2 pi freq.
 * op must be a field operator: idx = 1..3V (every voxel, in order), Mc = 1,
 * no Mr, an FP64 copy (VP_ERR_INVALID otherwise); then A(J) = Gram J - N(J).
 * J, Einc (nullable: zero), E: 3V complex128 following mem.
 */
int vp_field_total(vp_op op, const void* J, const void* Einc, double freq, void* E, int mem);

/*
 * Wavelength sweep by model order reduction (Bragg_MOR_script.m:1-143 and
 * example_Bragg_1500nm_to_1600nm.m:85-114).
 *
 * vp_op_create_lincomb: operator whose tensor is sum_i w[i] fN_i (w: k
 * interleaved complex128 weights, host), with the geometry, contrast and
 * plan of ops[0]; every source must share them (VP_ERR_INVALID otherwise).
 * With Lagrange weights this is lagrange_int(fN, S, s) (lagrange_int.m:1-13)
 * or lagrange_int_deriv (lagrange_int_deriv.m:1-20) applied to the tensors.
 *
 * vp_op_lincomb_into: the same combination written over the tensor of an
 * existing operator out (not one of ops; same geometry), allocating nothing:
 * fN_temp = lagrange_int(fN, S, s) of example_Bragg_1500nm_to_1600nm.m:104
 * once per swept wavelength.
 *
 * vp_vec_dots: out (k x m complex128, column-major, host) = X^H Y for k and m
 * device complex128 vectors of length n (the projections P'*A*P and P'*b).
 * vp_vec_lincomb: Y_j = sum_i X_i C(i, j) for C (k x m complex128,
 * column-major, host); Y must not alias X (the lift P*x).
 * vp_vec_absmax: out[i] = max_r |X_i[r]| (host doubles) for k device vectors:
 * the per-component max(max(max(abs(Einc)))) normalising the incident field
 * (example:90-91).
 * vp_vec_gather: Y[i] = alpha w[i] X[idx[i]] for i < n, idx 0-based device
 * int64 into X (length nx), w nullable (device complex128, 1 when NULL):
 * Vrhs = Gram tau3 Einc(idxS3) (example:93-95).  vp_vec_scatter: Y (length
 * ny) = 0, then Y[idx[i]] = X[i] (distinct indices): Je(idxS3) = x
 * (example:101-102).  An index out of range is VP_ERR_INVALID.
 * All vector entry points run on ctx's stream and return after it drains;
 * they share a context scratch buffer, so calls on one context serialise.
 */
int vp_op_create_lincomb(const vp_op* ops, const double* w, int k, vp_op* out);
int vp_op_lincomb_into(vp_op out, const vp_op* ops, const double* w, int k);
int vp_vec_dots(vp_ctx ctx, const void* const* X, int k, const void* const* Y, int m, int64_t n, double* out);
/*
 * vp_vec_dots_noalloc: the same X^H Y (same kernels, same results) with no
 * device allocation: its pointer tables, per-block partial sums and results
 * live in a 16 MiB array the library holds per GPU instead of the context's
 * scratch.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Limit: the scratch need,
 * 256-byte aligned 8k + 8m + 16 k m ceil(n / 1024) + 16 k m bytes, must fit
 * 16 MiB (e.g. k = m = 8 up to n = 16 M), else VP_ERR_UNSUPPORTED.  Every
 * call on one GPU shares its array: calls from threads of one process take
 * turns on a per-GPU lock, and the array keeps the last call's contents
 * (each call writes every entry it reads).
 */
int vp_vec_dots_noalloc(vp_ctx ctx, const void* const* X, int k, const void* const* Y, int m, int64_t n,
                        double* out);
int vp_vec_lincomb(vp_ctx ctx, const void* const* X, int k, const double* C, int m, int64_t n, void* const* Y);
int vp_vec_absmax(vp_ctx ctx, const void* const* X, int k, int64_t n, double* out);
int vp_vec_gather(vp_ctx ctx, const void* X, int64_t nx, const int64_t* idx, const void* w, double alpha_re,
                  double alpha_im, int64_t n, void* Y);
int vp_vec_scatter(vp_ctx ctx, const void* X, const int64_t* idx, int64_t n, void* Y, int64_t ny);

#ifdef __cplusplus
}
#endif
#endif /* VOXPHOTON_H */
