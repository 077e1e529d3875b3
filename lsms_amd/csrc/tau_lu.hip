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
// Pointers read from the solve descriptors are generic; casting them to the
// global address space lets the compiler emit global (not flat) loads/stores.
// (The host pass only type-checks kernel bodies, so it sees plain double2.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) double2 gdouble2;
#else
typedef double2 gdouble2;
#endif

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
// Complex 32x32x64 wave product on FP64 MFMA (v_mfma_f64_16x16x4f64), 3M form:
//   P1 = Ar Br, P2 = Ai Bi, P3 = (Ar+Ai)(Br+Bi);  re = P1 - P2, im = P3 - P1 - P2
// (three real MFMAs per complex 16x16x4 block instead of four).
// The MFMA is applied to the transposed product, D'(j,i) = sum_k B(k,j) A(i,k),
// so the accumulator of a lane covers consecutive rows i of C: register r of
// lane (li = lane%16, kq = lane/16) holds C(i0+li, j0+kq+4r), and every C and
// A access is a run of 16 consecutive rows (256 B).  Operands are read straight
// from L2/LDS into registers in MFMA layout (a-operand lane value B(k0+kq, j0+li),
// b-operand lane value A(i0+li, k0+kq)), one k-step ahead; no LDS staging and
// no barriers, so the waves of a CU run independently.
typedef d4 Acc[2][2];
#ifndef LSMSX_KUNROLL
#define LSMSX_KUNROLL 2
#endif

template <bool NEGA, class FA, class FB>
__device__ __forceinline__ void wave_gemm3m(FA fa, FB fb, Acc &p1, Acc &p2, Acc &p3) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  double2 a[2], b[2], an[2], bn[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    a[t] = fa(16 * t + li, kq);
    b[t] = fb(kq, 16 * t + li);
  }
#pragma unroll LSMSX_KUNROLL
  for (int k0 = 0; k0 < kTile; k0 += 4) {
    if (k0 + 4 < kTile) {
#pragma unroll
      for (int t = 0; t < 2; t++) {
        an[t] = fa(16 * t + li, k0 + 4 + kq);
        bn[t] = fb(k0 + 4 + kq, 16 * t + li);
      }
    }
    double as[2], bs[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      if (NEGA) {
        a[t].x = -a[t].x;
        a[t].y = -a[t].y;
      }
      as[t] = a[t].x + a[t].y;
      bs[t] = b[t].x + b[t].y;
    }
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
        p1[bj][bi] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[bj].x, a[bi].x, p1[bj][bi], 0, 0, 0);
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
        p2[bj][bi] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[bj].y, a[bi].y, p2[bj][bi], 0, 0, 0);
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
        p3[bj][bi] = __builtin_amdgcn_mfma_f64_16x16x4f64(bs[bj], as[bi], p3[bj][bi], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 2; t++) {
      a[t] = an[t];
      b[t] = bn[t];
    }
  }
}

__device__ __forceinline__ void zero_acc(Acc &p) {
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++) p[bj][bi] = d4{0.0, 0.0, 0.0, 0.0};
}

// C (wave's 32x32 block at Cg, leading dim ldc) = [C -] A B from the 3M sums.
template <bool SUBTRACT>
__device__ __forceinline__ void wave_store3m(double2 *__restrict__ Cg, long ldc, const Acc &p1,
                                             const Acc &p2, const Acc &p3) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        double2 *cp = Cg + (16 * bi + li) + ldc * (16 * bj + kq + 4 * r);
        double re = p1[bj][bi][r] - p2[bj][bi][r];
        double im = p3[bj][bi][r] - p1[bj][bi][r] - p2[bj][bi][r];
        if (SUBTRACT) {
          double2 c = *cp;
          *cp = make_double2(c.x - re, c.y - im);
        } else {
          *cp = make_double2(re, im);
        }
      }
}

// Panel of step s: x < nrest -> L_is = A_is U^-1 ; else U_sj = L^-1 P A_sj.
// The operand tile is staged in LDS (rows gathered by the pivot permutation
// for U_sj) before it is overwritten in place.
__global__ __launch_bounds__(256) void k_lu_panel(const SolveDesc *__restrict__ descs, int step) {
  const SolveDesc &d = descs[blockIdx.y];
  const int nrest = d.nt - 1 - step;
  const int x = blockIdx.x;
  if (nrest <= 0 || x >= 2 * nrest) return;
  __shared__ double2 T[kTile][kTile + 1];  // T[col][row]
  const int tid = threadIdx.x, w = tid >> 6;
  const int wr = 32 * (w & 1), wc = 32 * (w >> 1);
  const int c0 = step * kTile;
  double2 *M = d.M;
  const long ld = d.ld;
  const gdouble2 *__restrict__ inv = (const gdouble2 *)d.inv;
  const bool lower = x < nrest;
  double2 *tile = lower ? M + (long)(c0 + kTile * (1 + x)) + ld * c0
                        : M + (long)c0 + ld * (c0 + kTile * (1 + x - nrest));
  if (lower) {
    for (int e = tid; e < kTile * kTile; e += 256) {
      int r = e & 63, c = e >> 6;
      T[c][r] = tile[r + ld * c];
    }
  } else {
    const int *perm = d.perm;
    for (int e = tid; e < kTile * kTile; e += 256) {
      int r = e & 63, c = e >> 6;
      T[c][r] = tile[perm[r] + ld * c];
    }
  }
  __syncthreads();
  Acc p1, p2, p3;
  zero_acc(p1);
  zero_acc(p2);
  zero_acc(p3);
  if (lower) {
    // A(i,k) = T(wr+i, k); B(k,j) = U^-1(k, wc+j), upper incl. diagonal
    wave_gemm3m<false>([&](int i, int k) { return T[k][wr + i]; },
                [&](int k, int j) {
                  int jj = wc + j;
                  double2 v = make_double2(0.0, 0.0);
                  if (k <= jj) v = inv[k + kTile * jj];
                  return v;
                },
                p1, p2, p3);
  } else {
    // A(i,k) = L^-1(wr+i, k), unit lower; B(k,j) = (P A_sj)(k, wc+j)
    wave_gemm3m<false>([&](int i, int k) {
                  int ii = wr + i;
                  double2 v = make_double2(ii == k ? 1.0 : 0.0, 0.0);
                  if (ii > k) v = inv[ii + kTile * k];
                  return v;
                },
                [&](int k, int j) { return T[wc + j][k]; }, p1, p2, p3);
  }
  wave_store3m<false>(tile + wr + ld * wc, ld, p1, p2, p3);
}

// Trailing update of step s: A_ij -= L_is U_sj for all i, j > s.  One 64x64
// tile per workgroup, a 32x32 quarter per wave.  C enters the 3M sums at the
// start (q1 = Cre, q2 = 0, q3 = Cre + Cim, products accumulated with -A), so
// C' = (q1 - q2, q3 - q1 - q2) and no load sits in the epilogue.
__global__ __launch_bounds__(256) void k_lu_update(const SolveDesc *__restrict__ descs, int step) {
  const SolveDesc &d = descs[blockIdx.y];
  const int nrest = d.nt - 1 - step;
  const int x = blockIdx.x;
  if (nrest <= 0 || x >= nrest * nrest) return;
  const int ti = step + 1 + x % nrest, tj = step + 1 + x / nrest;
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  const int w = threadIdx.x >> 6;
  const int wr = 32 * (w & 1), wc = 32 * (w >> 1);
  const int c0 = step * kTile;
  double2 *M = d.M;
  const long ld = d.ld;
  const gdouble2 *__restrict__ A = (const gdouble2 *)(M + (long)(kTile * ti + wr) + ld * c0);
  const gdouble2 *__restrict__ B = (const gdouble2 *)(M + (long)c0 + ld * (kTile * tj + wc));
  gdouble2 *__restrict__ C = (gdouble2 *)(M + (long)(kTile * ti + wr) + ld * (kTile * tj + wc));
  Acc q1, q2, q3;
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++) {
      q2[bj][bi] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < 4; r++) {
        double2 c = C[(16 * bi + li) + ld * (16 * bj + kq + 4 * r)];  // global load
        q1[bj][bi][r] = c.x;
        q3[bj][bi][r] = c.x + c.y;
      }
    }
  wave_gemm3m<true>([&](int i, int k) { return A[i + ld * k]; },
                    [&](int k, int j) { return B[k + ld * j]; }, q1, q2, q3);
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        double re = q1[bj][bi][r] - q2[bj][bi][r];
        double im = q3[bj][bi][r] - q1[bj][bi][r] - q2[bj][bi][r];
        C[(16 * bi + li) + ld * (16 * bj + kq + 4 * r)] = make_double2(re, im);
      }
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
