// ir.hip -- iterative refinement of the mixed-precision tau00 solve.
//
// The FP32 factors P^T L U of the KKR matrix (tau_lu.hip, T = float) are a
// preconditioner; the answer is refined against the FP64 matrix M:
//   X_0 = 0;  R = B - M X (FP64, MFMA);  dX = (LU)^-1 R (FP32);  X += dX
// with B = [0; t0] in the solver layout (t0 in the central rows), until the
// central rows converge: ||dX_c||_F <= tol ||X_c||_F.  tau00 = X_c.
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
#include "lsmsx_device.hpp"

namespace lsmsx {

typedef double d4 __attribute__((ext_vector_type(4)));

#if defined(__HIP_DEVICE_COMPILE__)
#define IR_GLOBAL __attribute__((address_space(1)))
#else
#define IR_GLOBAL
#endif
typedef IR_GLOBAL double2 gd2;
typedef IR_GLOBAL float2 gf2;

// X = 0 and the first residual R = B - M*0 = B ([0; t0], t0 in the central
// rows), so the first sweep needs no pass over M.
__global__ __launch_bounds__(256) void k_ir_init(const SolveDesc *__restrict__ descs) {
  const SolveDesc &d = descs[blockIdx.x];
  gd2 *X = (gd2 *)d.X;
  gf2 *R = (gf2 *)d.R;
  const int Ns = d.Ns, kc = d.kc;
  const long n = (long)Ns * kc;
  for (long e = threadIdx.x; e < n; e += blockDim.x) {
    X[e] = make_double2(0.0, 0.0);
    const int row = (int)(e % Ns), col = (int)(e / Ns);
    float2 b = make_float2(0.f, 0.f);
    if (row >= Ns - kc) {
      const double2 t = d.t0[(row - (Ns - kc)) + (long)d.ldt * col];
      b = make_float2((float)t.x, (float)t.y);
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
__global__ __launch_bounds__(256) void k_ir_residual(const SolveDesc *__restrict__ descs) {
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
  gf2 *R = (gf2 *)d.R;
#pragma unroll
  for (int jb = 0; jb < 2; jb++) {
    if (jb < nb) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        double re = q[0][jb][r] - q[1][jb][r];
        double im = q[2][jb][r] - q[0][jb][r] - q[1][jb][r];
        R[(i0 + li) + (long)Ns * (16 * jb + kq + 4 * r)] = make_float2((float)re, (float)im);
      }
    }
  }
}

// Wave-level FP32 3M product for the refinement sweeps: acc(16 x 16nb) of
// rows i0.. += sum_k Bop(k, j) * Aop(i0 + i, k), transposed MFMA form as in the
// LU update (a-operand Bop(k0+kq, j0+li), b-operand Aop(i0+li, k0+kq)).
// Register r of lane (li, kq) holds output (row i0+li, column 16jb + 4kq + r).
typedef float f4 __attribute__((ext_vector_type(4)));
struct Acc3f {
  f4 p[3][2];
};

// The k range is consumed in batches of 8 k-steps whose operand loads are
// all issued before their MFMAs: one workgroup per solve leaves few waves per
// SIMD, so latency is hidden by loads in flight, not by occupancy.
template <class FA, class FB>
__device__ __forceinline__ void sweep_gemm(FA fa, FB fb, int k_begin, int k_end, int nb,
                                           Acc3f &q) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  constexpr int U = 8;
  int kb = k_begin;
  for (; kb + 4 * U <= k_end; kb += 4 * U) {
    float2 a[U], b[U][2];
#pragma unroll
    for (int u = 0; u < U; u++) {
      a[u] = fa(li, kb + 4 * u + kq);
      b[u][0] = fb(kb + 4 * u + kq, li);
      if (nb > 1) b[u][1] = fb(kb + 4 * u + kq, 16 + li);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const float as = a[u].x + a[u].y;
#pragma unroll
      for (int jb = 0; jb < 2; jb++) {
        if (jb < nb) {
          const float2 bb = b[u][jb];
          q.p[0][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bb.x, a[u].x, q.p[0][jb], 0, 0, 0);
          q.p[1][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bb.y, a[u].y, q.p[1][jb], 0, 0, 0);
          q.p[2][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bb.x + bb.y, as, q.p[2][jb], 0, 0, 0);
        }
      }
    }
  }
  for (int k0 = kb; k0 < k_end; k0 += 4) {
    float2 a = fa(li, k0 + kq);
    float as = a.x + a.y;
#pragma unroll
    for (int jb = 0; jb < 2; jb++) {
      if (jb < nb) {
        float2 b = fb(k0 + kq, 16 * jb + li);
        q.p[0][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(b.x, a.x, q.p[0][jb], 0, 0, 0);
        q.p[1][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(b.y, a.y, q.p[1][jb], 0, 0, 0);
        q.p[2][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(b.x + b.y, as, q.p[2][jb], 0, 0, 0);
      }
    }
  }
}

// 3M sums started from C: q1 = Cre, q2 = 0, q3 = Cre + Cim; with the
// products accumulated against -A this gives C - A B = (q1 - q2, q3 - q1 - q2).
__device__ __forceinline__ void acc_from(Acc3f &q, int jb, int r, float2 c) {
  q.p[0][jb][r] = c.x;
  q.p[1][jb][r] = 0.f;
  q.p[2][jb][r] = c.x + c.y;
}
__device__ __forceinline__ float2 acc_value(const Acc3f &q, int jb, int r) {
  return make_float2(q.p[0][jb][r] - q.p[1][jb][r], q.p[2][jb][r] - q.p[0][jb][r] - q.p[1][jb][r]);
}

// dX = (P^T L U)^-1 R in FP32 (in place in R), then X += dX and the
// convergence test on the central rows.  One workgroup per solve, the block
// rows of each step split over the 4 waves (16 rows each); the long sums
// sum_t L_st y_t and sum_t U_st x_t run on FP32 MFMA streaming the stored
// factors, the diagonal-tile inverses are applied from LDS.
__global__ __launch_bounds__(256) void k_ir_solve(const SolveDesc *__restrict__ descs, double tol) {
  const SolveDesc &d = descs[blockIdx.x];
  if (*d.conv) return;
  __shared__ float2 v[kTile][32 + 1];
  __shared__ double red[3][256];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, li = lane & 15, kq = lane >> 4;
  const int kc = d.kc, nt = d.nt, Ns = d.Ns, nb = kc / 16;
  const long ld = d.ld;
  const gf2 *__restrict__ W = (const gf2 *)d.A;
  gf2 *R = (gf2 *)d.R;
  // ---- forward: y_s = L_ss^-1 P_s (r_s - sum_{t<s} L_st y_t)
  for (int s = 0; s < nt; s++) {
    const int rs = s * kTile, i0 = rs + 16 * w;
    Acc3f q;
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) acc_from(q, jb, r, R[(i0 + li) + (long)Ns * (16 * jb + 4 * kq + r)]);
    sweep_gemm([&](int i, int k) {
                 float2 l = W[(i0 + i) + ld * k];
                 return make_float2(-l.x, -l.y);
               },
               [&](int k, int j) { return (float2)R[k + (long)Ns * j]; }, 0, rs, nb, q);
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) v[16 * w + li][16 * jb + 4 * kq + r] = acc_value(q, jb, r);
    __syncthreads();
    const int *perm = d.perm + s * kTile;
    const gf2 *inv = (const gf2 *)d.inv + s * d.inv_stride;
    // y = L^-1 (P v): unit lower L^-1, rows 16w.. of this wave
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        q.p[0][jb][r] = 0.f;
        q.p[1][jb][r] = 0.f;
        q.p[2][jb][r] = 0.f;
      }
    sweep_gemm([&](int i, int k) {
                 const int ii = 16 * w + i;
                 float2 l = make_float2(ii == k ? 1.f : 0.f, 0.f);
                 if (ii > k) l = inv[ii + kTile * k];
                 return l;
               },
               [&](int k, int j) { return v[perm[k]][j]; }, 0, kTile, nb, q);
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) R[(i0 + li) + (long)Ns * (16 * jb + 4 * kq + r)] = acc_value(q, jb, r);
    __threadfence_block();
    __syncthreads();
  }
  // ---- backward: x_s = U_ss^-1 (y_s - sum_{t>s} U_st x_t)
  for (int s = nt - 1; s >= 0; s--) {
    const int rs = s * kTile, i0 = rs + 16 * w;
    Acc3f q;
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) acc_from(q, jb, r, R[(i0 + li) + (long)Ns * (16 * jb + 4 * kq + r)]);
    sweep_gemm([&](int i, int k) {
                 float2 u = W[(i0 + i) + ld * k];
                 return make_float2(-u.x, -u.y);
               },
               [&](int k, int j) { return (float2)R[k + (long)Ns * j]; }, rs + kTile, Ns, nb, q);
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) v[16 * w + li][16 * jb + 4 * kq + r] = acc_value(q, jb, r);
    __syncthreads();
    const gf2 *inv = (const gf2 *)d.inv + s * d.inv_stride;
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        q.p[0][jb][r] = 0.f;
        q.p[1][jb][r] = 0.f;
        q.p[2][jb][r] = 0.f;
      }
    // x = U^-1 v: upper U^-1 incl. diagonal
    sweep_gemm([&](int i, int k) {
                 const int ii = 16 * w + i;
                 float2 u = make_float2(0.f, 0.f);
                 if (k >= ii) u = inv[ii + kTile * k];
                 return u;
               },
               [&](int k, int j) { return v[k][j]; }, 0, kTile, nb, q);
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (jb < nb) R[(i0 + li) + (long)Ns * (16 * jb + 4 * kq + r)] = acc_value(q, jb, r);
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
    float2 dx = R[e];
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

}  // namespace lsmsx
