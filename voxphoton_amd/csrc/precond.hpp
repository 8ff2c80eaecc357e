// Preconditioner objects and launch interfaces.
#pragma once
#include "op.hpp"

namespace vp {

struct PcGeom {
  int L, M, N;
  int64_t V, nS;        // voxels; unknown voxels (structured) or unknowns (generic)
  int structured;
  int level;            // 1 or 2
  int64_t n_outer;      // length of the vector the preconditioner acts on (embedding)
  int B, ld, tx, t2;    // pencil batch, LDS row stride, threads (x pass / plane)
  // 2-level blocks stored for kx < Hx, ky < Hy.  bsym = 1: Hx = L/2+1,
  // Hy = M/2+1 and B(L-kx, ky) = Sx B(kx, ky) Sx, B(kx, M-ky) = Sy B Sy with
  // Sx (Sy) negating the x (y) component; bsym = 0: Hx = L, Hy = M.
  int bsym, Hx, Hy;
  // 2-level apply split into passes when the fused (3N x M) plane exceeds
  // LDS: y DFT of Brow rows per workgroup (row stride ldrow), block products
  // over (kx, TK-column tiles), inverse y DFT.
  int psplit, Brow, ldrow, TK;
  // quadrant blocks (bsym) with 3N <= 255: y-row DFT / one workgroup per
  // stored block applying it to its four mirrored columns / inverse y-row DFT
  int pquad;
  // quadrant form over pencil-major spectra X[c][n][m][kx] with compile-time
  // x passes (length L) and row passes (length M); 0: plane-major X[kx][c][n][m]
  int pct;
  // columns of block loads in flight per chunk of the binary16 block product
  // (8, 16 or 32; 0: 16), timed when the FP32 preconditioner is made
  int qu = 0;
  // pct only: the block product runs over q-contiguous spectra Y[kx][m][q]
  // (q = c*N + n; each mirrored column one contiguous run) in work2, filled
  // from and returned to X by LDS-tiled transposes around it
  int qlay = 0;
  size_t lds_row, lds_tile;
  size_t lds1, lds2;
  FftPlan px, py;       // unpadded DFTs of length L and M
};

template <typename T>
struct PcArgs {
  const C<T>* x;
  C<T>* y;
  int64_t xstride, ystride;
  C<T>* work;           // 3 * L * M * N per problem
  C<T>* work2;          // 1-level: block product output; 2-level qlay: q-contiguous spectra
  int64_t wstride;
  const C<T>* const* blocks;   // device array: blocks of problem p
  const int* pos;
  const int* sub;       // box unknown -> outer vector index (nullptr: identity)
  const C<T>* twx;      // L entries
  const C<T>* twy;      // M entries
  const int* active;
  // FP32 apply of the quadrant form only: blocks stored as binary16 pairs
  // scaled by 1/hscale (nullptr: the blocks above)
  const uint32_t* const* hblocks;
  const float* hscale;
  // embedded apply: 0 copies x to y over the outer vector first (unknowns
  // outside the box pass through); 1 leaves y as it is (a later part of a
  // block-Jacobi composite, whose earlier parts already wrote y)
  int skip_copy;
};

// Times the binary16 block product of g's quadrant form with 8, 16 and 32
// columns of loads in flight (a few launches each on the preconditioner's own
// workspace) and returns the fastest window.
int tune_quadmul_window(const PcGeom& g, const PcArgs<float>& a, hipStream_t s);

template <typename T>
struct PrecPrec {
  bool present = false;
  DevBuf blocks;        // (kx, ky) major, column-major blocks, scaled by 1/(L*M)
  DevBuf blockptr;      // one device pointer to blocks
  DevBuf tw;            // L | M twiddles
  DevBuf work, work2;
  DevBuf xs, ys;        // staging for host calls
  // FP32 quadrant form: binary16 copy of the blocks (real, imaginary halves
  // of one 32-bit word), scaled by 1/hscale; pointer and scale as device
  // arrays of one entry for single-problem launches
  DevBuf hblocks, hblockptr, hscaleptr;
  float hscale = 1.f;
};

template <typename T>
void launch_prec(const PcGeom& g, const PcArgs<T>& a, int nprob, hipStream_t s);

// FP64 block construction (level_2_parallel_func.m semantics) and inversion.
// face_dev == nullptr: level_2_parallel_func semantics (A = -Mc*G, Gram*Mr0 =
// diag folded at the origin).  face_dev = device complex[N]: blocks are
// Gram*I - diag(face(n))*N2 with N2 the bare 2-level circulant of G (pass
// mc = -1, diag = 0).
// Blocks are built for kx < kxn, ky < kyn at (kx * kyn + ky).
void build_level2_blocks(const double2* G, int L, int M, int N, double2 mc, double diag, double2* blk,
                         int kxn, int kyn, hipStream_t s, const double2* face_dev = nullptr, double gram = 0.0);
// 1-level (3MN x 3MN per kx, all L) blocks, the same conventions; face_dev =
// device complex[M*N] for the face-averaged form Gram*I - diag(face3)*circ_N.
void build_level1_blocks(const double2* G, int L, int M, int N, double2 mc, double diag, double2* blk,
                         hipStream_t s, const double2* face_dev = nullptr, double gram = 0.0);
// Inverts blocks of any size: the device Gauss-Jordan kernels up to 256 rows,
// rocSOLVER's batched LU (getrf / getri, loaded on first use) above.
void invert_blocks_any(double2* blk, int64_t nblocks, int n, hipStream_t s);
// True when the odd components of a device Toeplitz tensor (xy, xz, yz)
// vanish on their zero-offset planes (to 1e-12 of max |G|), which makes the
// Chan circulants parity symmetric.
bool toeplitz_parity_zero(const double2* G, int L, int M, int N, hipStream_t s);
void invert_blocks(double2* blk, int64_t nblocks, int n, hipStream_t s);
// binary16 copy of the FP32 blocks (P.hblocks, P.hscale) for the quadrant form
void make_half_blocks(PrecPrec<float>& P, hipStream_t s);

}  // namespace vp

struct vp_prec_s {
  vp_ctx ctx = nullptr;
  vp::PcGeom g{};
  int64_t nbox = 0;          // unknowns inside the box (|idx|)
  double scale32 = 1.0;      // FP32 copy inverts A / scale32
  vp::DevBuf pos;
  vp::DevBuf sub;            // embedding (optional)
  // content hashes of the voxel map and the embedding: batched launches
  // share one map, so only preconditioners with equal maps batch together
  uint64_t pos_hash = 0, sub_hash = 0;
  std::vector<int> sub_host;   // the embedding (0-based outer positions), host copy
  // block-Jacobi composite (prec_fun_disk_idx): the parts, each embedded in
  // the same outer vector on disjoint unknowns; not owned (they outlive the
  // composite).  Empty for a preconditioner with its own blocks.
  std::vector<vp_prec_s*> parts;
  vp::PrecPrec<double> p64;
  vp::PrecPrec<float> p32;
};

namespace vp {
// FNV-1a over the entries of an index map.
inline uint64_t map_hash(const std::vector<int>& v) {
  uint64_t h = 1469598103934665603ull;
  for (int e : v) {
    h ^= uint64_t(uint32_t(e));
    h *= 1099511628211ull;
  }
  return h ^ uint64_t(v.size());
}

// The preconditioners whose blocks an apply runs: the parts of a composite,
// else the preconditioner itself.
inline std::vector<vp_prec_s*> prec_leaves(vp_prec_s* p) {
  if (!p) return {};
  if (!p->parts.empty()) return p->parts;
  return {p};
}

template <typename T> PrecPrec<T>& pprec_of(vp_prec_s* p);
template <> inline PrecPrec<double>& pprec_of<double>(vp_prec_s* p) { return p->p64; }
template <> inline PrecPrec<float>& pprec_of<float>(vp_prec_s* p) { return p->p32; }

template <typename T>
inline PcArgs<T> prec_args(vp_prec_s* p, const C<T>* x, C<T>* y) {
  PrecPrec<T>& P = pprec_of<T>(p);
  PcArgs<T> a{};
  a.x = x;
  a.y = y;
  a.xstride = 0;
  a.ystride = 0;
  a.work = P.work.template as<C<T>>();
  a.work2 = P.work2.template as<C<T>>();
  a.wstride = 0;
  a.blocks = P.blockptr.template as<const C<T>*>();
  a.pos = p->pos.as<int>();
  a.sub = p->sub.p ? p->sub.as<int>() : nullptr;
  a.twx = P.tw.template as<C<T>>();
  a.twy = P.tw.template as<C<T>>() + p->g.L;
  a.active = nullptr;
  a.hblocks = P.hblocks.p ? P.hblockptr.template as<const uint32_t*>() : nullptr;
  a.hscale = P.hblocks.p ? P.hscaleptr.template as<float>() : nullptr;
  return a;
}
}  // namespace vp
