// Internal structures shared by the VoxPhoton device library translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "common.hpp"
#include "fft.hpp"

namespace vp {

// ----- error handling -------------------------------------------------------
enum Status { OK = 0, ERR_INVALID = -1, ERR_OOM = -2, ERR_DEVICE = -3, ERR_UNSUPPORTED = -4 };
void set_error(const std::string& msg);
struct VpError {
  int code;
  std::string msg;
};
#define VP_HIP(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::vp::VpError{e_ == hipErrorOutOfMemory ? ::vp::ERR_OOM : ::vp::ERR_DEVICE, \
                          std::string(#expr) + ": " + hipGetErrorString(e_)};          \
  } while (0)
#define VP_REQUIRE(cond, code, msg)                   \
  do {                                                \
    if (!(cond)) throw ::vp::VpError{(code), (msg)}; \
  } while (0)

// ----- geometry of one operator ----------------------------------------------
// Grid L x M x N (column-major voxels, l fastest), FFT sizes Lp x Mp x Np.
// Each FFT dimension is split into R residue classes k = R*a + p (p < R);
// a class is a DFT of length S = P / R of the folded, twiddled input.
struct Dim {
  int D;      // grid extent
  int P;      // padded FFT length
  int R;      // number of residue classes
  int S;      // class DFT length (P / R)
  FftPlan plan;
};

struct MatvecPlan {
  // 0: fused plane kernel over residue classes; 1: separate y / z / y^-1 passes
  int variant;
  // split variant: y passes (By pencils of Mp per workgroup), z pass (Tz ky columns)
  int By, ldy, ty, Tz, tz;
  int fsym;      // split variant: fN stored as its parity octant
  int zr1, zr2;  // split variant: register z-column kernel Np = zr1 * zr2 (zr1 = 0: LDS z pass)
  int zra, zrb, ztpk;  // split variant: fused two-stage z kernel Np = zra * zrb, ztpk threads per ky column (0: off)
  size_t lds_zf;       // its LDS bytes
  int yct;             // split variant: compile-time y pass of length Mp (0: generic y pass)
  int xct;             // split variant: compile-time x passes of class length S over pencil-major
                       // x spectra X[c][n][m][kx] (0: generic x passes, plane-major X[kx][c][n][m])
  FftPlan zf, zi; // split variant: z plans of the forward (odd radices first) and inverse transforms
  size_t lds_y, lds_z;
  // K1/K3 (x pencils)
  int B;         // pencils per workgroup
  int ldsx;      // row stride of a pencil in LDS (elements)
  int tx;        // threads per workgroup
  int ept3;      // accumulator elements per thread in K3
  // K2 (yz planes)
  int t2;        // threads per workgroup
  int ept2;      // accumulator elements per thread
  size_t lds1, lds2;  // dynamic LDS bytes
};

// L, M, N are the extents of the unknowns' bounding box (l0, m0, n0 its
// offset in the Lg x Mg x . grid): the circulant convolution is shift
// invariant, so every pass runs on the box and only the voxel index uses the
// full grid.  V is the full grid's voxel count.
struct Geometry {
  int L, M, N;
  int l0, m0, n0;
  int64_t Lg, Mg;
  int64_t V;
  Dim x, y, z;
  int64_t nS;        // unknown voxels (structured) or unknown entries (generic)
  int structured;    // 1: idx = [idxS; V+idxS; 2V+idxS] with sorted idxS
  __host__ __device__ int64_t voxel(int l, int m, int n) const {
    return (l0 + l) + Lg * ((m0 + m) + Mg * int64_t(n0 + n));
  }
};

// Static device pointers of one problem of a (possibly batched) launch.  The
// vectors themselves are kernel arguments: problem p reads x + p * vstride
// and writes y + p * vstride.
template <typename T>
struct ProbPtrs {
  C<T>* work;        // 3 * Lp * M * N spectrum workspace
  C<T>* work2;       // 3 * Lp * Mp * N (split variant only)
  const C<T>* fN;    // tensor in class layout, 6 * Lp * Mp * Np, prescaled
  const int* pos;    // voxel -> unknown map (V or 3V entries, -1 = not an unknown)
  const C<T>* mc;    // contrast per unknown voxel (nS) / per unknown (generic n)
  const C<T>* mr;    // Mr per unknown voxel; nullptr means Mr == 1
  T gram;            // identity weight in this precision
};

template <typename T>
struct VecArgs {
  const C<T>* x;
  C<T>* y;
  int64_t xstride;      // elements between consecutive problems' input vectors
  int64_t ystride;      // ... and output vectors
  const int* active;    // per-problem flag (nullptr: all active)
};

// Per-kernel event timing (prof.cpp)
enum KernelId {
  KID_XFWD = 0, KID_PLANE, KID_XINV, KID_PC_XFWD, KID_PC_PLANE, KID_PC_XINV, KID_YFWD, KID_ZPASS, KID_YINV,
  KID_PC_YROWS, KID_PC_BLOCK,
  kNumKernelIds
};
const char* kernel_id_name(int kid);
void prof_begin(hipStream_t s, int kid, void** token);
void prof_end(hipStream_t s, void* token);
void prof_enable(hipStream_t s, bool on);
bool prof_enabled_on(hipStream_t s);
void prof_stats(hipStream_t s, int kid, long long* launches, double* ms, bool reset);
struct ProfScope {
  hipStream_t s;
  void* tok;
  ProfScope(hipStream_t s_, int kid) : s(s_), tok(nullptr) { prof_begin(s, kid, &tok); }
  ~ProfScope() { prof_end(s, tok); }
};

// Host-side planning helpers (plan.cpp)
bool factorize(int n, bool fp64, FftPlan& out, int tws, bool odd_first = false);
// Fused two-stage z kernel: ky columns per workgroup, radix pair for Np and
// threads per column (plan.cpp).
constexpr int kZfTile = 16;
bool zfused_pair(int Np, bool fp64, int& ra, int& rb);
// Threads per ky column: at least RA (stage B gives one butterfly per
// thread), at most 256 / kZfTile; chosen by a VALU cost model of the three
// stages (A and B' loop over 3*RB items, B runs once).  A compile-time
// constant of the kernel and the launch width on the host.
__host__ __device__ constexpr int zfused_threads(int ra, int rb) {
  auto dftc = [](int r) {
    return r == 2 ? 4 : r == 3 ? 12 : r == 4 ? 16 : r == 5 ? 32 : r == 7 ? 64 : r == 8 ? 56 : r == 11 ? 130 : 144;
  };
  const int items = 3 * rb;
  const long long ca = dftc(ra) + 3 * ra, cb = 3 * (2 * dftc(rb) + 5 * rb) + 40 * rb;
  int best = ra;
  long long bestc = -1;
  for (int tpk = ra; tpk <= 256 / kZfTile; ++tpk) {
    const long long passes = (items + tpk - 1) / tpk;
    const long long c = tpk * (2 * passes * ca + cb);
    if (bestc < 0 || c < bestc) { bestc = c; best = tpk; }
  }
  return best;
}
// Compile-time y pass: pencils per workgroup for length n and element size
// (two ping-pong buffers of n x (NB + 1) plus the twiddle table within
// 40 KB), and the radix list instantiated for n (plan.cpp; false: none).
__host__ __device__ constexpr int yct_pencils(int n, int elem) {
  int nb = 16;
  while (nb > 2 && (2 * n * (nb + 1) + n) * elem > 40 * 1024) nb /= 2;
  return nb;
}
__host__ __device__ constexpr int yct_lds(int n, int elem) { return (2 * n * (yct_pencils(n, elem) + 1) + n) * elem; }
// Lengths with a compile-time y pass and their radix lists (both precisions).
#define VP_YCT_LIST(X)                                                                                  \
  X(12, 3, 4) X(16, 4, 4) X(20, 4, 5) X(24, 3, 8) X(32, 4, 8) X(40, 5, 8) X(48, 3, 4, 4) X(64, 8, 8)     \
  X(96, 3, 4, 8) X(100, 4, 5, 5) X(120, 3, 5, 8) X(128, 2, 8, 8) X(200, 5, 5, 8) X(256, 4, 8, 8)
bool yct_supported(int n);
// Class lengths S = Lp/R with compile-time x passes:
// X(S, forward threads, (forward radices), inverse threads, (inverse radices)).
// The forward pass gathers through the voxel map in its first stage, so a
// first stage with many butterflies (small radix) keeps more lanes loading;
// the inverse reads contiguous data and ends in LDS.
#define VP_UNPAREN(...) __VA_ARGS__
#define VP_XCT32_LIST(X)                                                                                \
  X(3600, 768, (5, 5, 9, 16), 512, (16, 9, 5, 5)) X(1800, 512, (8, 9, 5, 5), 512, (8, 9, 5, 5))           \
  X(1200, 512, (16, 3, 5, 5), 512, (16, 3, 5, 5)) X(900, 256, (4, 9, 5, 5), 256, (4, 9, 5, 5))             \
  X(450, 256, (2, 9, 5, 5), 256, (2, 9, 5, 5)) X(400, 256, (16, 5, 5), 256, (16, 5, 5))                   \
  X(360, 256, (8, 9, 5), 256, (8, 9, 5)) X(200, 256, (8, 5, 5), 256, (8, 5, 5)) X(100, 256, (4, 5, 5), 256, (4, 5, 5))
#define VP_XCT64_LIST(X)                                                                                \
  X(3600, 512, (8, 2, 9, 5, 5), 512, (8, 2, 9, 5, 5)) X(1800, 512, (8, 9, 5, 5), 512, (8, 9, 5, 5))       \
  X(1200, 512, (8, 2, 3, 5, 5), 512, (8, 2, 3, 5, 5)) X(900, 256, (4, 9, 5, 5), 256, (4, 9, 5, 5))         \
  X(450, 256, (2, 9, 5, 5), 256, (2, 9, 5, 5)) X(400, 256, (4, 4, 5, 5), 256, (4, 4, 5, 5))               \
  X(360, 256, (8, 9, 5), 256, (8, 9, 5)) X(200, 256, (8, 5, 5), 256, (8, 5, 5)) X(100, 256, (4, 5, 5), 256, (4, 5, 5))
// Preconditioner row lengths (unpadded circulant axis) with compile-time
// row passes, and their radix lists.
#define VP_PCROW_LIST(X)                                                                                \
  X(11, 11) X(22, 2, 11) X(10, 2, 5) X(12, 3, 4) X(20, 4, 5) X(24, 3, 8) X(44, 4, 11) X(48, 3, 4, 4)       \
  X(16, 4, 4) X(32, 4, 8)
bool pcrow_supported(int n);
// Threads of the compile-time x pass for class length S (0: none).
int xct_threads(int S, bool fp64);
int max_lds_bytes(int device);

template <typename T>
void launch_matvec(const Geometry& g, const MatvecPlan& mp, const ProbPtrs<T>* d_probs, VecArgs<T> va,
                   int nprob, int transp, const C<T>* twx, const C<T>* twy, const C<T>* twz, hipStream_t s);

bool make_matvec_plan(const Geometry& g, bool fp64, int device, MatvecPlan& out, std::string& why, int variant);

}  // namespace vp
