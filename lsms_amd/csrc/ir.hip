// ir.hip -- iterative refinement of tau00 against the FP64 KKR matrix.
//
// The factors P^T L U of a copy of the KKR matrix (tau_lu.hip) are a
// preconditioner; the answer is refined against the FP64 matrix M:
//   X_0 = 0;  R = B - M X (FP64, MFMA);  dX = (LU)^-1 R;  X += dX
// with B = [0; t0] in the solver layout (t0 in the central rows), until the
// central rows converge: ||dX_c||_F <= tol ||X_c||_F.  tau00 = X_c.
// Two instantiations of the correction type T:
//   float   mixed precision: complex FP32 factors and sweeps;
//   double  FP64 factors and sweeps, used when the tile-local pivots of a
//           plain FP64 factorization come out small, to recover full FP64
//           accuracy (residuals against the untouched M).
// The reference's only mixed-precision routes are LAPACK zcgesv
// (linearSolvers_CPU.cpp:148-174, which never fills its right-hand side) and
// cusolverDnIRSXgesv (linearSolvers_CUDA.cu:425-470); both refine a full
// solve the same way.
//
// Forward and backward sweeps use the block structure of the tile-pivoted
// factorization: for tile row s,
//   y_s = L_ss^-1 P_s (r_s - sum_{t<s} L_st y_t)      (L_st stored left of the diagonal tile)
//   x_s = U_ss^-1 (y_s - sum_{t>s} U_st x_t)
// with the diagonal-tile inverses and permutations kept by the factorization.
#include "lu_common.hpp"

namespace lsmsx {

typedef GCx<double> gd2;

// X = 0 and the first residual R = B - M*0 = B ([0; t0], t0 in the central
// rows), so the first sweep needs no pass over M.
template <class T>
__global__ __launch_bounds__(256) void k_ir_init(const SolveDesc *__restrict__ descs) {
  typedef CxT<T> V;
  const SolveDesc &d = descs[blockIdx.x];
  gd2 *X = (gd2 *)d.X;
  GCx<T> *R = (GCx<T> *)d.R;
  const int Ns = d.Ns, kc = d.kc;
  const long n = (long)Ns * kc;
  for (long e = threadIdx.x; e < n; e += blockDim.x) {
    X[e] = make_double2(0.0, 0.0);
    const int row = (int)(e % Ns), col = (int)(e / Ns);
    V b;
    b.x = T(0);
    b.y = T(0);
    if (row >= Ns - kc) {
      const double2 t = d.t0[(row - (Ns - kc)) + (long)d.ldt * col];
      b.x = (T)t.x;
      b.y = (T)t.y;
    }
    R[e] = b;
  }
  if (threadIdx.x == 0) {
    *d.conv = 0;
    *d.iters = 0;
    *d.dxn = 0.0;
    *d.dxn_rho2 = 0.0;
  }
}

// R = B - M X for one 64-row tile (blockIdx.x) of one solve (blockIdx.y);
// wave w handles rows 16w.. of the tile, all kc right-hand sides (16 or 32).
// FP64 MFMA in the 3M, transposed form of the LU update:
//   a-operand X(k0+kq, j0+li), b-operand M(i0+li, k0+kq).
// The residual is stored in the correction type T.
template <class T>
__global__ __launch_bounds__(256) void k_ir_residual(const SolveDesc *__restrict__ descs) {
  typedef CxT<T> V;
  const SolveDesc &d = descs[blockIdx.y];
  if (*d.conv) return;
  const int tile = blockIdx.x;
  if (tile >= d.nt) return;
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  const int w = threadIdx.x >> 6;
  const int i0 = tile * kTile + 16 * w;
  const long ld = d.ld;
  const int Ns = d.Ns, kc = d.kc, nb = kc / 16;
  const gd2 *__restrict__ M = (const gd2 *)d.M + i0;
  const gd2 *__restrict__ X = (const gd2 *)d.X;
  d4 q[3][2];
  // accumulators start from B: rows >= Ns - kc hold t0
#pragma unroll
  for (int jb = 0; jb < 2; jb++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int row = i0 + li, col = 16 * jb + kq + 4 * r;
      double2 b = make_double2(0.0, 0.0);
      if (jb < nb && row >= Ns - kc) b = d.t0[(row - (Ns - kc)) + (long)d.ldt * col];
      q[0][jb][r] = b.x;
      q[1][jb][r] = 0.0;
      q[2][jb][r] = b.x + b.y;
    }
  for (int k0 = 0; k0 < Ns; k0 += 4) {
    double2 m = M[li + ld * (k0 + kq)];
    double mr = -m.x, mi = -m.y, ms = mr + mi;
#pragma unroll
    for (int jb = 0; jb < 2; jb++) {
      if (jb < nb) {
        double2 x = X[(k0 + kq) + (long)Ns * (16 * jb + li)];
        q[0][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, mr, q[0][jb], 0, 0, 0);
        q[1][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, mi, q[1][jb], 0, 0, 0);
        q[2][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x + x.y, ms, q[2][jb], 0, 0, 0);
      }
    }
  }
  GCx<T> *R = (GCx<T> *)d.R;
#pragma unroll
  for (int jb = 0; jb < 2; jb++) {
    if (jb < nb) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        V o;
        o.x = (T)(q[0][jb][r] - q[1][jb][r]);
        o.y = (T)(q[2][jb][r] - q[0][jb][r] - q[1][jb][r]);
        R[(i0 + li) + (long)Ns * (16 * jb + kq + 4 * r)] = o;
      }
    }
  }
}

// Wave-level 3M product for the refinement sweeps: acc(16 x 16nb) of rows
// i0.. += sum_k Bop(k, j) * Aop(i0 + i, k), transposed MFMA form as in the LU
// update (a-operand Bop(k0+kq, j0+li), b-operand Aop(i0+li, k0+kq)).
// Register r of lane (li, kq) holds output (row i0+li, column 16jb +
// acc_row<T>(kq, r)).
template <class T>
struct AccIR {
  Acc4<T> p[3][2];
};

// The k range is consumed in batches of 8 k-steps whose operand loads are
// all issued before their MFMAs: one workgroup per solve leaves few waves per
// SIMD, so latency is hidden by loads in flight, not by occupancy.
template <class T, class FA, class FB>
__device__ __forceinline__ void sweep_gemm(FA fa, FB fb, int k_begin, int k_end, int nb,
                                           AccIR<T> &q) {
  typedef CxT<T> V;
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  constexpr int U = 8;
  int kb = k_begin;
  for (; kb + 4 * U <= k_end; kb += 4 * U) {
    V a[U], b[U][2];
#pragma unroll
    for (int u = 0; u < U; u++) {
      a[u] = fa(li, kb + 4 * u + kq);
      b[u][0] = fb(kb + 4 * u + kq, li);
      if (nb > 1) b[u][1] = fb(kb + 4 * u + kq, 16 + li);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const T as = a[u].x + a[u].y;
#pragma unroll
      for (int jb = 0; jb < 2; jb++) {
        if (jb < nb) {
          const V bb = b[u][jb];
          q.p[0][jb] = mfma4(bb.x, a[u].x, q.p[0][jb]);
          q.p[1][jb] = mfma4(bb.y, a[u].y, q.p[1][jb]);
          q.p[2][jb] = mfma4(bb.x + bb.y, as, q.p[2][jb]);
        }
      }
    }
  }
  for (int k0 = kb; k0 < k_end; k0 += 4) {
    V a = fa(li, k0 + kq);
    T as = a.x + a.y;
#pragma unroll
    for (int jb = 0; jb < 2; jb++) {
      if (jb < nb) {
        V b = fb(k0 + kq, 16 * jb + li);
        q.p[0][jb] = mfma4(b.x, a.x, q.p[0][jb]);
        q.p[1][jb] = mfma4(b.y, a.y, q.p[1][jb]);
        q.p[2][jb] = mfma4(b.x + b.y, as, q.p[2][jb]);
      }
    }
  }
}

// 3M sums started from C: q1 = Cre, q2 = 0, q3 = Cre + Cim; with the
// products accumulated against -A this gives C - A B = (q1 - q2, q3 - q1 - q2).
template <class T>
__device__ __forceinline__ void acc_from(AccIR<T> &q, int jb, int r, CxT<T> c) {
  q.p[0][jb][r] = c.x;
  q.p[1][jb][r] = T(0);
  q.p[2][jb][r] = c.x + c.y;
}
template <class T>
__device__ __forceinline__ CxT<T> acc_value(const AccIR<T> &q, int jb, int r) {
  CxT<T> v;
  v.x = q.p[0][jb][r] - q.p[1][jb][r];
  v.y = q.p[2][jb][r] - q.p[0][jb][r] - q.p[1][jb][r];
  return v;
}
template <class T>
__device__ __forceinline__ void acc_zero(AccIR<T> &q) {
#pragma unroll
  for (int jb = 0; jb < 2; jb++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      q.p[0][jb][r] = T(0);
      q.p[1][jb][r] = T(0);
      q.p[2][jb][r] = T(0);
    }
}

// dX = (P^T L U)^-1 R in the correction type (in place in R), then X += dX
// and the convergence test on the central rows.  One workgroup per solve, the
// block rows of each step split over the 4 waves (16 rows each); the long
// sums sum_t L_st y_t and sum_t U_st x_t run on MFMA streaming the stored
// factors, the diagonal-tile inverses are applied from LDS.
template <class T>
__global__ __launch_bounds__(256) void k_ir_solve(const SolveDesc *__restrict__ descs, double tol) {
  typedef CxT<T> V;
  const SolveDesc &d = descs[blockIdx.x];
  if (*d.conv) return;
  __shared__ V v[kTile][32 + 1];
  __shared__ double red[3][256];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, li = lane & 15, kq = lane >> 4;
  const int kc = d.kc, nt = d.nt, Ns = d.Ns, nb = kc / 16;
  const long ld = d.ld;
  const GCx<T> *__restrict__ W = (const GCx<T> *)d.A;
  GCx<T> *R = (GCx<T> *)d.R;
  auto col_of = [&](int jb, int r) { return 16 * jb + acc_row<T>(kq, r); };
  // ---- forward: y_s = L_ss^-1 P_s (r_s - sum_{t<s} L_st y_t)
  for (int s = 0; s < nt; s++) {
    const int rs = s * kTile, i0 = rs + 16 * w;
    AccIR<T> q;
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) acc_from<T>(q, jb, r, R[(i0 + li) + (long)Ns * col_of(jb, r)]);
    sweep_gemm<T>([&](int i, int k) {
                    V l = W[(i0 + i) + ld * k];
                    l.x = -l.x;
                    l.y = -l.y;
                    return l;
                  },
                  [&](int k, int j) { return (V)R[k + (long)Ns * j]; }, 0, rs, nb, q);
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) v[16 * w + li][col_of(jb, r)] = acc_value<T>(q, jb, r);
    __syncthreads();
    const int *perm = d.perm + s * kTile;
    const GCx<T> *inv = (const GCx<T> *)d.inv + s * d.inv_stride;
    // y = L^-1 (P v): unit lower L^-1, rows 16w.. of this wave
    acc_zero<T>(q);
    sweep_gemm<T>([&](int i, int k) {
                    const int ii = 16 * w + i;
                    V l;
                    l.x = ii == k ? T(1) : T(0);
                    l.y = T(0);
                    if (ii > k) l = inv[ii + kTile * k];
                    return l;
                  },
                  [&](int k, int j) { return v[perm[k]][j]; }, 0, kTile, nb, q);
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) R[(i0 + li) + (long)Ns * col_of(jb, r)] = acc_value<T>(q, jb, r);
    __threadfence_block();
    __syncthreads();
  }
  // ---- backward: x_s = U_ss^-1 (y_s - sum_{t>s} U_st x_t)
  for (int s = nt - 1; s >= 0; s--) {
    const int rs = s * kTile, i0 = rs + 16 * w;
    AccIR<T> q;
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) acc_from<T>(q, jb, r, R[(i0 + li) + (long)Ns * col_of(jb, r)]);
    sweep_gemm<T>([&](int i, int k) {
                    V u = W[(i0 + i) + ld * k];
                    u.x = -u.x;
                    u.y = -u.y;
                    return u;
                  },
                  [&](int k, int j) { return (V)R[k + (long)Ns * j]; }, rs + kTile, Ns, nb, q);
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) v[16 * w + li][col_of(jb, r)] = acc_value<T>(q, jb, r);
    __syncthreads();
    const GCx<T> *inv = (const GCx<T> *)d.inv + s * d.inv_stride;
    acc_zero<T>(q);
    // x = U^-1 v: upper U^-1 incl. diagonal
    sweep_gemm<T>([&](int i, int k) {
                    const int ii = 16 * w + i;
                    V u;
                    u.x = T(0);
                    u.y = T(0);
                    if (k >= ii) u = inv[ii + kTile * k];
                    return u;
                  },
                  [&](int k, int j) { return v[k][j]; }, 0, kTile, nb, q);
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) R[(i0 + li) + (long)Ns * col_of(jb, r)] = acc_value<T>(q, jb, r);
    __threadfence_block();
    __syncthreads();
  }
  // ---- X += dX; convergence on the central rows.  The next error is about
  // rho ||dX|| with rho = ||dX_k|| / ||dX_{k-1}|| the observed contraction;
  // converge when that estimate (and ||dX|| itself, to 1e-8) is below tol.
  gd2 *X = (gd2 *)d.X;
  double dn = 0.0, xn = 0.0;
  for (long e = tid; e < (long)Ns * kc; e += 256) {
    const int row = (int)(e % Ns);
    const V dx = R[e];
    double2 x = X[e];
    x.x += dx.x;
    x.y += dx.y;
    X[e] = x;
    if (row >= Ns - kc) {
      dn += (double)dx.x * dx.x + (double)dx.y * dx.y;
      xn += x.x * x.x + x.y * x.y;
    }
  }
  red[0][tid] = dn;
  red[1][tid] = xn;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) {
      red[0][tid] += red[0][tid + off];
      red[1][tid] += red[1][tid + off];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double dn2 = red[0][0], xn2 = red[1][0];
    const int k = *d.iters;
    const double prev = *d.dxn;  // previous ||dX_c||^2
    bool ok = false;
    if (k >= 1 && prev > 0.0) {
      // the larger of the last two observed contractions, so that one
      // unusually small ratio does not end the refinement early
      const double rho2 = fmax(dn2 / prev, *d.dxn_rho2);
      ok = (rho2 < 1.0) && (rho2 * dn2 <= tol * tol * xn2) && (dn2 <= 1e-16 * xn2);
      *d.dxn_rho2 = dn2 / prev;
    }
    if (dn2 <= tol * tol * xn2) ok = true;
    *d.iters = k + 1;
    *d.dxn = dn2;
    if (ok) *d.conv = 1;
  }
}

// tau00 = central rows of X.
__global__ __launch_bounds__(256) void k_ir_extract(const SolveDesc *__restrict__ descs) {
  const SolveDesc &d = descs[blockIdx.x];
  const int kc = d.kc, Ns = d.Ns;
  for (int e = threadIdx.x; e < kc * kc; e += blockDim.x) {
    int r = e % kc, c = e / kc;
    d.tau00[r + kc * c] = d.X[(Ns - kc + r) + (long)Ns * c];
  }
}

__global__ void k_count_conv(const SolveDesc *__restrict__ descs, int n, int *count) {
  __shared__ int s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  int local = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) local += *descs[i].conv;
  atomicAdd(&s, local);
  __syncthreads();
  if (threadIdx.x == 0) *count = s;
}

// Copy the FP64 matrix M into the factorization buffer A (FP64), one 64-column
// strip per workgroup, for the FP64-refined route.
__global__ __launch_bounds__(256) void k_copy_m(const SolveDesc *__restrict__ descs) {
  const SolveDesc &d = descs[blockIdx.y];
  const int strip = blockIdx.x;
  if (strip >= d.nt) return;
  const long ld = d.ld, c0 = (long)strip * kTile;
  const gd2 *__restrict__ M = (const gd2 *)d.M;
  gd2 *A = (gd2 *)d.A;
  for (long e = threadIdx.x; e < (long)kTile * d.Ns; e += 256) {
    const long c = c0 + e / d.Ns, r = e % d.Ns;
    A[r + ld * c] = M[r + ld * c];
  }
}

template __global__ void k_ir_init<float>(const SolveDesc *);
template __global__ void k_ir_init<double>(const SolveDesc *);
template __global__ void k_ir_residual<float>(const SolveDesc *);
template __global__ void k_ir_residual<double>(const SolveDesc *);
template __global__ void k_ir_solve<float>(const SolveDesc *, double);
template __global__ void k_ir_solve<double>(const SolveDesc *, double);

}  // namespace lsmsx
