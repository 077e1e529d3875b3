// tau_lu.hip -- batched blocked complex LU of KKR matrices on MI355X,
// returning only the central block tau00 = S^-1 t0, where S is the Schur
// complement of the non-central sites.
//
// Reference behaviour replaced: solveTau00zgetrf (linearSolvers_CPU.cpp:98-123:
// LU of M, solve [t0;0], keep the top kkrsz_ns rows) and the block-inverse
// route (zblock_lu_cpp.cpp:19-86 + tau_inv_postproc.f:33-57: eliminate the
// non-central blocks with pivoting inside diagonal blocks, then
// (1-delta)^-1 t0).  Here the matrix is in the solver layout (central site
// last, identity padding first) and the elimination is right-looking over
// 64-wide tiles with partial pivoting inside each 64x64 diagonal tile, the
// block-LU pivoting scheme of the reference's block-inverse solvers.
//
// Per step s (tile column c0 = 64 s), batched over all solves:
//   k_lu_diag    factor A_ss = P L U in LDS; store L^-1 and U^-1 (one tile)
//   k_lu_panel   L_is = A_is U^-1 (tiles below), U_sj = L^-1 P A_sj (right)
//   k_lu_update  A_ij -= L_is U_sj                     <- FP64 MFMA, dominant
// then k_lu_finish eliminates the last tile's non-central part, factors the
// Schur complement S and solves S tau00 = t0.
#include "lsmsx_device.hpp"

namespace lsmsx {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline double cabs1(double2 z) { return fabs(z.x) + fabs(z.y); }
__device__ inline double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 crecip(double2 b) {
  double s = 1.0 / (b.x * b.x + b.y * b.y);
  return make_double2(b.x * s, -b.y * s);
}

// Wave-wide argmax of v (ties -> smallest index); all lanes get the result.
__device__ inline void wave_argmax(double &v, int &idx) {
  for (int off = 32; off >= 1; off >>= 1) {
    double ov = __shfl_xor(v, off, 64);
    int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

// In-LDS LU with partial pivoting of the leading `ncols` columns of a 64x64
// tile A[col][row], pivot rows restricted to [j, prow_end).  Full-row swaps
// and full-width updates (all 64 columns, rows below j).  Records the swap
// sequence in piv[] and the permutation in perm[].  256 threads.
__device__ void tile_lu_lds(double2 (*A)[kTile + 1], int ncols, int prow_end, int *piv,
                            int *perm, int *sinfo, double *sminpiv, int col_base) {
  const int tid = threadIdx.x;
  __shared__ int s_p;
  for (int j = 0; j < ncols; j++) {
    if (tid < 64) {
      double v = (tid >= j && tid < prow_end) ? cabs1(A[j][tid]) : -1.0;
      int idx = tid;
      wave_argmax(v, idx);
      if (tid == 0) {
        s_p = idx;
        piv[j] = idx;
        int t = perm[j];
        perm[j] = perm[idx];
        perm[idx] = t;
        double a = sqrt(A[j][idx].x * A[j][idx].x + A[j][idx].y * A[j][idx].y);
        if (a < *sminpiv) *sminpiv = a;
        if (v == 0.0 && *sinfo == 0) *sinfo = col_base + j + 1;
      }
    }
    __syncthreads();
    const int p = s_p;
    if (p != j && tid < 64) {
      double2 t = A[tid][j];
      A[tid][j] = A[tid][p];
      A[tid][p] = t;
    }
    __syncthreads();
    const double2 a = A[j][j];
    const bool nz = (a.x != 0.0 || a.y != 0.0);
    if (nz && tid > j && tid < 64) A[j][tid] = cmul2(A[j][tid], crecip(a));
    __syncthreads();
    if (nz) {
      for (int e = tid; e < 64 * 64; e += 256) {
        int r = e & 63, c = e >> 6;
        if (r > j && c > j) {
          double2 l = A[j][r], u = A[c][j];
          A[c][r].x -= l.x * u.x - l.y * u.y;
          A[c][r].y -= l.x * u.y + l.y * u.x;
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Diagonal tile of step s: factor, write back L\U, store L^-1 (strict lower)
// and U^-1 (upper incl. diagonal) packed in one tile, and the permutation.
__global__ __launch_bounds__(256) void k_lu_diag(const SolveDesc *__restrict__ descs, int step) {
  const SolveDesc &d = descs[blockIdx.x];
  if (step >= d.nt - 1) return;
  __shared__ double2 A[kTile][kTile + 1];
  __shared__ int piv[kTile], perm[kTile];
  __shared__ int sinfo;
  __shared__ double sminpiv;
  const int tid = threadIdx.x;
  const int c0 = step * kTile;
  double2 *M = d.M;
  const long ld = d.ld;
  for (int e = tid; e < kTile * kTile; e += 256) {
    int r = e & 63, c = e >> 6;
    A[c][r] = M[(c0 + r) + ld * (c0 + c)];
  }
  if (tid < kTile) perm[tid] = tid;
  if (tid == 0) {
    sinfo = 0;
    sminpiv = 1e300;
  }
  __syncthreads();
  tile_lu_lds(A, kTile, kTile, piv, perm, &sinfo, &sminpiv, c0);
  // write back the factors
  for (int e = tid; e < kTile * kTile; e += 256) {
    int r = e & 63, c = e >> 6;
    M[(c0 + r) + ld * (c0 + c)] = A[c][r];
  }
  if (tid < kTile) d.perm[tid] = perm[tid];
  if (tid == 0) {
    if (sinfo != 0 && *d.info == 0) *d.info = sinfo;
    if (sminpiv < *d.minpiv) *d.minpiv = sminpiv;
  }
  __syncthreads();
  // In-place inversion: unit-lower L (threads 0..63 own rows) and upper U
  // (threads 64..127 own rows).  L^-1: for j = 62..0, X[r][j] = -(L[r][j] +
  // sum_{j<k<r} X[r][k] L[k][j]) for r > j.  U^-1: for j = 0..63,
  // X[j][j] = 1/U[j][j], X[r][j] = -X[j][j] sum_{r<=k<j} X[r][k] U[k][j].
  for (int jj = 0; jj < kTile; jj++) {
    double2 nl = make_double2(0.0, 0.0), nu = make_double2(0.0, 0.0);
    const int jl = kTile - 1 - jj;  // lower: columns from the right
    const int ju = jj;              // upper: columns from the left
    if (tid < 64) {
      int r = tid;
      if (r > jl) {
        double2 s = A[jl][r];
        for (int k = jl + 1; k < r; k++) {
          double2 x = A[k][r], l = A[jl][k];
          s.x += x.x * l.x - x.y * l.y;
          s.y += x.x * l.y + x.y * l.x;
        }
        nl = make_double2(-s.x, -s.y);
      }
    } else if (tid < 128) {
      int r = tid - 64;
      double2 ujj = A[ju][ju];
      bool nz = (ujj.x != 0.0 || ujj.y != 0.0);
      double2 xjj = nz ? crecip(ujj) : make_double2(0.0, 0.0);
      if (r < ju) {
        double2 s = make_double2(0.0, 0.0);
        for (int k = r; k < ju; k++) {
          double2 x = A[k][r], u = A[ju][k];
          s.x += x.x * u.x - x.y * u.y;
          s.y += x.x * u.y + x.y * u.x;
        }
        double2 v = cmul2(s, xjj);
        nu = make_double2(-v.x, -v.y);
      } else if (r == ju) {
        nu = xjj;
      }
    }
    __syncthreads();
    if (tid < 64) {
      if (tid > jl) A[jl][tid] = nl;
    } else if (tid < 128) {
      int r = tid - 64;
      if (r <= ju) A[ju][r] = nu;
    }
    __syncthreads();
  }
  double2 *inv = d.inv;
  for (int e = tid; e < kTile * kTile; e += 256) {
    int r = e & 63, c = e >> 6;
    inv[r + kTile * c] = A[c][r];
  }
}

// ---------------------------------------------------------------------------
// 64x64x64 complex tile product on FP64 MFMA (v_mfma_f64_16x16x4f64).
// Operands staged in LDS as split re/im planes, k-major:
//   As[k][i] (A(i,k)),  Bs[k][j] (B(k,j)), in two K-chunks of 32.
// Wave w computes rows 32*(w&1).., cols 32*(w>>1).. as 2x2 16x16 blocks with
// 4 real MFMAs per complex block (re += ar br - ai bi, im += ar bi + ai br).
// MFMA 16x16x4 f64 operand layout: a = A[lane%16][lane/16],
// b = B[lane/16][lane%16]; accumulator register r of a lane holds
// D[lane/16 + 4r][lane%16] (rows interleaved across the 4 lane groups).
constexpr int kKc = 32;
constexpr int kBs = kTile + 1;  // padded stride of Bs rows

struct GemmSmem {
  double are[kKc][kTile], aim[kKc][kTile];
  double bre[kKc][kBs], bim[kKc][kBs];
};

// Operand sources.  amode: 0 global tile, 1 packed L^-1 (unit lower).
// bmode: 0 global tile, 1 packed U^-1 (upper), 2 global tile with rows gathered by perm.
template <int AMODE, int BMODE, bool NEGA>
__device__ inline void stage_chunk(GemmSmem &s, const double2 *__restrict__ Ag, long lda,
                                   const double2 *__restrict__ Bg, long ldb, const int *perm,
                                   int k0) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int it = 0; it < (kTile * kKc) / 256; it++) {
    int e = tid + 256 * it;
    int i = e & 63, k = e >> 6;  // A(i, k0+k)
    double2 v;
    if (AMODE == 0) {
      v = Ag[i + lda * (k0 + k)];
    } else {
      int kk = k0 + k;
      v = (i > kk) ? Ag[i + kTile * kk] : make_double2(i == kk ? 1.0 : 0.0, 0.0);
    }
    if (NEGA) {
      v.x = -v.x;
      v.y = -v.y;
    }
    s.are[k][i] = v.x;
    s.aim[k][i] = v.y;
  }
#pragma unroll
  for (int it = 0; it < (kTile * kKc) / 256; it++) {
    int e = tid + 256 * it;
    int k = e & 31, jcol = e >> 5;  // B(k0+k, jcol)
    int kk = k0 + k;
    double2 v;
    if (BMODE == 0) {
      v = Bg[kk + ldb * jcol];
    } else if (BMODE == 1) {
      v = (kk <= jcol) ? Bg[kk + kTile * jcol] : make_double2(0.0, 0.0);
    } else {
      v = Bg[perm[kk] + ldb * jcol];
    }
    s.bre[k][jcol] = v.x;
    s.bim[k][jcol] = v.y;
  }
}

__device__ inline void mma_chunk(const GemmSmem &s, d4 (&cre)[2][2], d4 (&cim)[2][2], int wr,
                                 int wc, int lane) {
  const int li = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < kKc; kk += 4) {
    double ar[2], ai[2], br[2], bi[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
      ar[b] = s.are[kk + kq][wr + 16 * b + li];
      ai[b] = s.aim[kk + kq][wr + 16 * b + li];
      br[b] = s.bre[kk + kq][wc + 16 * b + li];
      bi[b] = s.bim[kk + kq][wc + 16 * b + li];
    }
#pragma unroll
    for (int ib = 0; ib < 2; ib++)
#pragma unroll
      for (int jb = 0; jb < 2; jb++) {
        cre[ib][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[ib], br[jb], cre[ib][jb], 0, 0, 0);
        cre[ib][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[ib], bi[jb], cre[ib][jb], 0, 0, 0);
        cim[ib][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[ib], bi[jb], cim[ib][jb], 0, 0, 0);
        cim[ib][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[ib], br[jb], cim[ib][jb], 0, 0, 0);
      }
  }
}

// C (global, 64x64) = [C +] (+-A) B.
template <int AMODE, int BMODE, bool ACCUM>
__device__ void tile_gemm(GemmSmem &s, double2 *__restrict__ Cg, long ldc,
                          const double2 *__restrict__ Ag, long lda,
                          const double2 *__restrict__ Bg, long ldb, const int *perm) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int wr = 32 * (w & 1), wc = 32 * (w >> 1);
  const int li = lane & 15, kq = lane >> 4;
  d4 cre[2][2], cim[2][2];
#pragma unroll
  for (int ib = 0; ib < 2; ib++)
#pragma unroll
    for (int jb = 0; jb < 2; jb++) {
      if (ACCUM) {
        const double2 *cp = Cg + (wr + 16 * ib + kq) + ldc * (wc + 16 * jb + li);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          double2 v = cp[4 * r];
          cre[ib][jb][r] = v.x;
          cim[ib][jb][r] = v.y;
        }
      } else {
        cre[ib][jb] = d4{0.0, 0.0, 0.0, 0.0};
        cim[ib][jb] = d4{0.0, 0.0, 0.0, 0.0};
      }
    }
  for (int k0 = 0; k0 < kTile; k0 += kKc) {
    __syncthreads();
    stage_chunk<AMODE, BMODE, ACCUM>(s, Ag, lda, Bg, ldb, perm, k0);
    __syncthreads();
    mma_chunk(s, cre, cim, wr, wc, lane);
  }
#pragma unroll
  for (int ib = 0; ib < 2; ib++)
#pragma unroll
    for (int jb = 0; jb < 2; jb++) {
      double2 *cp = Cg + (wr + 16 * ib + kq) + ldc * (wc + 16 * jb + li);
#pragma unroll
      for (int r = 0; r < 4; r++) cp[4 * r] = make_double2(cre[ib][jb][r], cim[ib][jb][r]);
    }
}

// Panel of step s: x < nrest -> L_is = A_is U^-1 ; else U_sj = L^-1 P A_sj.
// The tile is read completely into LDS before it is overwritten in place.
__global__ __launch_bounds__(256) void k_lu_panel(const SolveDesc *__restrict__ descs, int step) {
  const SolveDesc &d = descs[blockIdx.y];
  const int nrest = d.nt - 1 - step;
  const int x = blockIdx.x;
  if (nrest <= 0 || x >= 2 * nrest) return;
  __shared__ GemmSmem s;
  __shared__ double2 T[kTile][kTile];  // copy of the operand tile (column-major)
  __shared__ int perm[kTile];
  const int tid = threadIdx.x;
  const int c0 = step * kTile;
  double2 *M = d.M;
  const long ld = d.ld;
  if (tid < kTile) perm[tid] = d.perm[tid];
  double2 *tile;
  if (x < nrest) tile = M + (long)(c0 + kTile * (1 + x)) + ld * c0;                 // A_is
  else tile = M + (long)c0 + ld * (c0 + kTile * (1 + x - nrest));                   // A_sj
  for (int e = tid; e < kTile * kTile; e += 256) {
    int r = e & 63, c = e >> 6;
    T[c][r] = tile[r + ld * c];
  }
  __syncthreads();
  if (x < nrest) tile_gemm<0, 1, false>(s, tile, ld, &T[0][0], kTile, d.inv, kTile, perm);
  else tile_gemm<1, 2, false>(s, tile, ld, d.inv, kTile, &T[0][0], kTile, perm);
}

// Trailing update of step s: A_ij -= L_is U_sj for all i, j > s.
__global__ __launch_bounds__(256) void k_lu_update(const SolveDesc *__restrict__ descs, int step) {
  const SolveDesc &d = descs[blockIdx.y];
  const int nrest = d.nt - 1 - step;
  const int x = blockIdx.x;
  if (nrest <= 0 || x >= nrest * nrest) return;
  __shared__ GemmSmem s;
  const int ti = step + 1 + x % nrest, tj = step + 1 + x / nrest;
  const int c0 = step * kTile;
  double2 *M = d.M;
  const long ld = d.ld;
  tile_gemm<0, 0, true>(s, M + (long)kTile * ti + ld * (kTile * tj), ld,
                        M + (long)kTile * ti + ld * c0, ld, M + (long)c0 + ld * (kTile * tj), ld,
                        nullptr);
}

// Last tile: eliminate its non-central part (pivot rows restricted to it),
// factor the Schur complement S (central rows), and solve S tau00 = t0.
__global__ __launch_bounds__(256) void k_lu_finish(const SolveDesc *__restrict__ descs) {
  const SolveDesc &d = descs[blockIdx.x];
  __shared__ double2 A[kTile][kTile + 1];
  __shared__ double2 X[kTile][kTile + 1];
  __shared__ int piv[kTile], perm[kTile];
  __shared__ int sinfo;
  __shared__ double sminpiv;
  const int tid = threadIdx.x;
  const int c0 = d.Ns - kTile;
  const int kc = d.kc, nd = kTile - kc;
  double2 *M = d.M;
  const long ld = d.ld;
  for (int e = tid; e < kTile * kTile; e += 256) {
    int r = e & 63, c = e >> 6;
    A[c][r] = M[(c0 + r) + ld * (c0 + c)];
  }
  if (tid < kTile) perm[tid] = tid;
  if (tid == 0) {
    sinfo = 0;
    sminpiv = 1e300;
  }
  __syncthreads();
  tile_lu_lds(A, nd, nd, piv, perm, &sinfo, &sminpiv, c0);
  // The trailing kc x kc block is S; continue the elimination on it with
  // pivoting among the central rows.
  for (int j = nd; j < kTile; j++) {
    __shared__ int s_p;
    if (tid < 64) {
      double v = (tid >= j) ? cabs1(A[j][tid]) : -1.0;
      int idx = tid;
      wave_argmax(v, idx);
      if (tid == 0) {
        s_p = idx;
        piv[j] = idx;
        double a = sqrt(A[j][idx].x * A[j][idx].x + A[j][idx].y * A[j][idx].y);
        if (a < sminpiv) sminpiv = a;
        if (v == 0.0 && sinfo == 0) sinfo = c0 + j + 1;
      }
    }
    __syncthreads();
    const int p = s_p;
    if (p != j && tid >= nd && tid < 64) {
      double2 t = A[tid][j];
      A[tid][j] = A[tid][p];
      A[tid][p] = t;
    }
    __syncthreads();
    const double2 a = A[j][j];
    const bool nz = (a.x != 0.0 || a.y != 0.0);
    if (nz && tid > j && tid < 64) A[j][tid] = cmul2(A[j][tid], crecip(a));
    __syncthreads();
    if (nz) {
      for (int e = tid; e < 64 * 64; e += 256) {
        int r = e & 63, c = e >> 6;
        if (r > j && c > j) {
          double2 l = A[j][r], u = A[c][j];
          A[c][r].x -= l.x * u.x - l.y * u.y;
          A[c][r].y -= l.x * u.y + l.y * u.x;
        }
      }
    }
    __syncthreads();
  }
  // X = P_S t0, then unit-lower and upper solves; one thread per RHS column.
  for (int e = tid; e < kc * kc; e += 256) {
    int r = e % kc, c = e / kc;
    X[c][r] = d.t0[r + (long)d.ldt * c];
  }
  __syncthreads();
  if (tid < kc) {
    const int c = tid;
    for (int j = nd; j < kTile; j++) {
      int p = piv[j];
      if (p != j) {
        double2 t = X[c][j - nd];
        X[c][j - nd] = X[c][p - nd];
        X[c][p - nd] = t;
      }
    }
    for (int j = 0; j < kc; j++) {
      double2 xj = X[c][j];
      for (int r = j + 1; r < kc; r++) {
        double2 l = A[nd + j][nd + r];
        X[c][r].x -= l.x * xj.x - l.y * xj.y;
        X[c][r].y -= l.x * xj.y + l.y * xj.x;
      }
    }
    for (int j = kc - 1; j >= 0; j--) {
      double2 u = A[nd + j][nd + j];
      double2 xj = (u.x != 0.0 || u.y != 0.0) ? cmul2(X[c][j], crecip(u)) : X[c][j];
      X[c][j] = xj;
      for (int r = 0; r < j; r++) {
        double2 uu = A[nd + j][nd + r];
        X[c][r].x -= uu.x * xj.x - uu.y * xj.y;
        X[c][r].y -= uu.x * xj.y + uu.y * xj.x;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < kc * kc; e += 256) {
    int r = e % kc, c = e / kc;
    d.tau00[r + kc * c] = X[c][r];
  }
  if (tid == 0) {
    if (sinfo != 0 && *d.info == 0) *d.info = sinfo;
    if (sminpiv < *d.minpiv) *d.minpiv = sminpiv;
  }
}

}  // namespace lsmsx
