// lsmsx_device.hpp -- shared device-side types of the MI355X tau00 path.
//
// Matrix layout used by the solver ("solver layout"): an Ns x Ns column-major
// complex matrix, Ns = roundup(N, 64).  The first pad = Ns - N rows and
// columns hold an identity block (eliminated trivially), then the LIZ sites
// 1..numLIZ-1 in LIZ order, then the central site 0 last, so the central
// kkrsz_ns x kkrsz_ns block is the trailing corner of the last 64-tile and
// tau00 comes from the Schur complement of everything before it.
#pragma once
#include <hip/hip_runtime.h>

namespace lsmsx {

constexpr int kTile = 64;     // LU tile / panel width
constexpr int kMaxLmax = 3;   // tables and per-site blocks sized for lmax <= 3
constexpr int kMaxKkr = (kMaxLmax + 1) * (kMaxLmax + 1);
constexpr int kPlaneHalves = 4 * kTile * kTile;  // one tile's four FP16 planes
constexpr int kIrFirstSweep = -1000;  // k_ir_step's step value for the first forward sweep
// k_ir_sweep (one launch per refinement sweep, kc = 16): waves per workgroup
// and the largest number of 64-row tiles per solve it takes
constexpr int kSweepWaves = 8;
constexpr int kSweepMaxTiles = 128;
constexpr unsigned kSweepSpinLimit = 1u << 24;  // bound of one k_ir_sweep flag wait
constexpr int kInfoSweepStall = -2;             // per-solve info when that bound is hit
// per-solve info of a solve whose FP64 refinement did not converge: its
// tau00 is the unrefined FP64 result
constexpr int kInfoRefineUnconverged = -3;  // LSMSX_INFO_REFINE_UNCONVERGED

// Element (r, c) of a tile-major matrix with nt tiles per tile column: every
// 64 x 64 tile is one contiguous run (column-major inside), tiles ordered by
// tile column, then tile row.  A tile column of complex FP32 values is then
// one 512-byte run and a whole tile one 32 KiB run.
static_assert(kTile == 64, "tiled_index assumes 64 x 64 tiles");
__host__ __device__ inline long tiled_index(int r, int c, int nt) {
  return ((long)(c >> 6) * nt + (r >> 6)) * (kTile * kTile) + (c & 63) * kTile + (r & 63);
}

// One solve (local atom, spin, energy slot) of a batch.
struct SolveDesc {
  // LIZ of the atom (device pointers)
  const double *pos;      // 3 x numLIZ, relative to site 0
  const int *lmax;        // numLIZ
  const int *store;       // numLIZ: t-matrix store column of each site
  const int *site_off;    // numLIZ: first row/col of each site's block in M
  int numLIZ;
  int n_spin_cant;
  int kkrsz;              // (maxlmax+1)^2 of the central atom
  int kc;                 // kkrsz_ns: central block size (= number of right-hand sides)
  double prel_re, prel_im;
  long t_off;             // offset of this energy slot (and collinear spin block) in a tmatStore column
  int ldt;                // leading dimension of one t block (n_spin_cant * kkrsz)
  // matrices
  double2 *M;             // FP64 KKR matrix (assembly target; residuals of the mixed path)
  void *A;                // matrix the LU kernels factor: M itself (FP64) or a float2 copy
  int ld;                 // leading dimension of M (and of A when not tiled)
  int tiled;              // 1: A is tile-major (tiled_index), 0: column-major with ld
  int Ns;                 // matrix dimension in this layout
  int pad;                // identity rows/cols at the front
  int nt;                 // Ns / kTile (solver layout)
  int step0;              // global step at which this solve's elimination starts
  int group;              // elimination steps per trailing update (2, or 4 with split planes)
  // factorization state
  void *inv;              // per diagonal tile: L^-1 (strict lower) and U^-1 (upper), kTile^2
  long inv_stride;        // elements between the tiles of consecutive steps (0: one reused tile)
  int *perm;              // nt x kTile row permutations of the diagonal tiles
  int *info;              // first singular pivot (1-based, solver order) or 0
  double *minpiv;         // smallest |pivot| seen
  double2 *tau00;         // kc x kc output
  const double2 *t0;      // central t-matrix (kc x kc, leading dim ldt)
  // iterative refinement (mixed precision)
  double2 *X;             // Ns x kc solution, FP64 (refined paths only)
  void *R;                // Ns x kc residual / correction, in the factor precision
  float2 *w32;            // FP32 copy of M written by the assembly (mixed path) or null
  int *conv;              // 1 once converged
  int *iters;             // refinement sweeps used
  double *dxn;            // ||dX_c||^2 of the previous sweep
  double *dxn_rho2;       // the contraction (||dX_c|| ratio)^2 observed at the previous sweep
  double *nrm;            // ||dX_c||^2, ||X_c||^2 of the current sweep
  // split-FP16 trailing updates (mixed path): the FP32 panels of the steps
  // of a group as hi/lo FP16 planes with one power-of-2 scale per tile
  void *hp;               // (step % group) x (L, U) x nt tiles x 4 planes x 64 x 64 halves
  int *he;                // (step % group) x (L, U) x nt tile exponents
  // the current diagonal tile's inverses as split planes for the fused panel:
  // L^-1 (unit lower, as an L tile) then U^-1 (upper, as a U tile), and
  // their two exponents; written by k_lu_diag<float> when set
  void *ihp;              // 2 tiles x 4 planes x 64 x 64 halves
  int *ihe;               // 2
  // explicit inverse (full tau, Lloyd term, adjoint of tau00)
  int full_lu;            // 1: the last tile is factored by k_lu_diag like every other tile
  double2 *Minv;          // Ns x Ns (leading dim ld): M^-1 in the solver layout
  double dprel_re, dprel_im;  // d prel / dE
  double2 *gx;            // Ns x kc: the rows of site i hold (G X)_i, X = M^-1 [0; t0]
  double2 *lpart;         // numLIZ per-row-site parts of Tr(M^-1 dM/dE)
  // mixed path with the structure constants stored once per site pair (no
  // FP64 M): the residual is formed as B - X + T (G X)
  double2 *gh;            // n(n-1)/2 blocks G_pq, p < q solver positions (gh_index), 16 x 16
                          // column-major (rows: site p), so G_qp = D G_pq D, D = (-1)^l
  double2 *ybuf;          // n site blocks: the G X rows of each band's own band row
  double2 *gpart;         // band slots (c, b <= c) x 16 sites: rows of band c from band b
  const double2 *tmat;    // t-matrix store and its leading dimension (site t of the residual)
  long ldT;
};

// Outputs of one k_assemble launch: the FP64 M; M and its tile-major FP32
// copy W; or the structure constants once per site pair (FP64) and W
// (collinear only)
constexpr int kAsmM = 0;
constexpr int kAsmMW = 1;
constexpr int kAsmGW = 2;
// the structure constants alone (once per group of solves that share the
// LIZ geometry and the energy; k_assemble_from_g then writes every solve's
// blocks from them)
constexpr int kAsmG = 3;

// G-half residual: sites per band, waves per workgroup and sites per wave
constexpr int kGBand = 16;
constexpr int kGWaves = 8;
constexpr int kGSites = kGBand / kGWaves;

// Block of the solver-position pair p < q among n sites (row-major upper triangle).
__host__ __device__ inline long gh_index(int p, int q, int n) {
  return (long)p * (2 * n - p - 1) / 2 + (q - p - 1);
}
// Solver position of LIZ site s (site 0, the central one, last).
__host__ __device__ inline int solver_pos(int s, int n) { return s ? s - 1 : n - 1; }
// LIZ site at solver position p.
__host__ __device__ inline int site_at_pos(int p, int n) { return p == n - 1 ? 0 : p + 1; }

// one potential of a single-site batch: 1-based jmt, jws, iend = max(jmt+11, jws)
// as in semrel_c.f, and the 0-based four-point window of the matching
// interpolation at r_sph (cinterp.f: interval ip1-1..ip1, outside points i1, i2)
struct SingleSitePot {
  int jmt, jws, iend;
  int ip1, i1, i2;
  double r_sph;
};

// one potential of a muffin-tin Green-function batch: 1-based jmt, jws, and
// the 0-based window of the interpolation at xs = sqrt(r_sph) on the mesh
// x_0 = 0, x_k = sqrt(r_k) (k = 1..jmt): interval ip1-1..ip1, outside points i1, i2
constexpr int kGreenSolves = 8;  // solves per k_green_mt workgroup (8 lanes each)
struct GreenPot {
  int jmt, jws;
  int ip1, i1, i2;
  double xs;
};

}  // namespace lsmsx
