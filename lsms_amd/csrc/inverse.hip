// inverse.hip -- explicit inverse of KKR matrices from their tile-pivoted LU,
// and the kernels of the full-tau and adjoint routes built on it.
//
// Reference behaviour replaced: solveTauFullzgetrf (linearSolvers_CPU.cpp:125-144:
// zgetrf + zgetrs with the N x N block-diagonal t as right-hand side, tau =
// M^-1 T, used by Kubo/CurrentMatrix.cpp:184-190).  The reference has no
// Lloyd term or adjoint; here the same inverse gives
//   Tr(M^-1 dM/dE) = d ln det M / dE          (k_lloyd, kkr_assembly.hip)
//   d tau00 = sum_i R_i dt_i C_i + R_0 dt_0   (k_tau_grad)
// with R_i = (M^-1)_{0i} the central block row and C_i = (G M^-1)_{i0} t_0.
//
// Factorization (tau_lu.hip with full_lu = 1 and every diagonal-tile inverse
// kept): M = Lt Ut, Lt block lower with diagonal tiles P_s^T L_ss, Ut block
// upper with diagonal tiles U_ss; Lt_ts (t > s) and Ut_ts (t < s) sit in M.
// M^-1 = Ut^-1 Lt^-1 is formed in place in one Ns x Ns buffer Y:
//   forward,  s = 0..nt-1:  Y_ss = L_ss^-1 P_s,  Y_sc = L_ss^-1 P_s Y_sc (c < s)
//                           Y_tc -= Lt_ts Y_sc   (t > s, c <= s; Lt^-1 is block lower)
//   backward, s = nt-1..0:  Y_sc = U_ss^-1 Y_sc  (all c)
//                           Y_tc -= Ut_ts Y_sc   (t < s, all c)
// Work: Ns^3/6 (forward) + Ns^3/2 (backward) complex MACs on FP64 MFMA, 3M form.
#include "lu_common.hpp"

namespace lsmsx {

typedef CxT<double> Z;
typedef GCx<double> GZ;

// Store a wave's 32x32 3M accumulator block (rows wr.., cols wc.. of C).
__device__ __forceinline__ void store_acc(GZ *C, long ld, const Acc3<double> &q) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        Z o;
        o.x = q.p[0][bj][bi][r] - q.p[1][bj][bi][r];
        o.y = q.p[2][bj][bi][r] - q.p[0][bj][bi][r] - q.p[1][bj][bi][r];
        C[(16 * bi + li) + ld * (16 * bj + acc_row<double>(kq, r))] = o;
      }
}

// C -= A B for one 64x64 tile, C entering the 3M sums (as k_lu_update).
__device__ __forceinline__ void tile_update(const GZ *__restrict__ A, const GZ *__restrict__ B,
                                            GZ *__restrict__ C, long ld) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  const int w = threadIdx.x >> 6;
  const int wr = 32 * (w & 1), wc = 32 * (w >> 1);
  A += wr;
  B += ld * wc;
  C += wr + ld * wc;
  Acc3<double> q;
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        Z c = C[(16 * bi + li) + ld * (16 * bj + acc_row<double>(kq, r))];
        q.p[0][bj][bi][r] = c.x;
        q.p[1][bj][bi][r] = 0.0;
        q.p[2][bj][bi][r] = c.x + c.y;
      }
  wave_gemm3m<double, true>([&](int i, int k) { return A[i + ld * k]; },
                            [&](int k, int j) { return B[k + ld * j]; }, q);
  store_acc(C, ld, q);
}

// Fill Y with zeros: one 64-column strip per workgroup.
__global__ __launch_bounds__(256) void k_inv_zero(const SolveDesc *__restrict__ descs) {
  const SolveDesc &d = descs[blockIdx.y];
  const int strip = blockIdx.x;
  if (strip >= d.nt) return;
  GZ *Y = (GZ *)d.Minv + d.ld * (long)(strip * kTile);
  const long n = (long)kTile * d.ld;
  for (long e = threadIdx.x; e < n; e += 256) Y[e] = Z{0.0, 0.0};
}

// Forward step s, diagonal part: Y_sc = L_ss^-1 P_s Y_sc for c < s and
// Y_ss = L_ss^-1 P_s (column tile c = blockIdx.x <= s).
__global__ __launch_bounds__(256) void k_inv_fwd_panel(const SolveDesc *__restrict__ descs,
                                                       int step) {
  const SolveDesc &d = descs[blockIdx.y];
  const int c = blockIdx.x;
  if (step >= d.nt || c > step) return;
  __shared__ Z Ts[kTile][kTile + 1];  // Ts[col][row]
  const int tid = threadIdx.x, w = tid >> 6;
  const int wr = 32 * (w & 1), wc = 32 * (w >> 1);
  const long ld = d.ld;
  GZ *tile = (GZ *)d.Minv + (long)(step * kTile) + ld * (c * kTile);
  const int *perm = d.perm + step * kTile;
  const GZ *__restrict__ inv = (const GZ *)d.inv + step * d.inv_stride;
  for (int e = tid; e < kTile * kTile; e += 256) {
    int r = e & 63, col = e >> 6;
    Z v;
    if (c == step) v = Z{perm[r] == col ? 1.0 : 0.0, 0.0};
    else v = tile[perm[r] + ld * col];
    Ts[col][r] = v;
  }
  __syncthreads();
  Acc3<double> q;
  zero_acc(q);
  // A(i,k) = L^-1(wr+i, k), unit lower; B(k,j) = (P Y_sc)(k, wc+j)
  wave_gemm3m<double, false>([&](int i, int k) {
                               int ii = wr + i;
                               Z v = Z{ii == k ? 1.0 : 0.0, 0.0};
                               if (ii > k) v = inv[ii + kTile * k];
                               return v;
                             },
                             [&](int k, int j) { return Ts[wc + j][k]; }, q);
  store_acc(tile + wr + ld * wc, ld, q);
}

// Forward step s, update: Y_tc -= Lt_ts Y_sc for t > s, c <= s.
__global__ __launch_bounds__(256) void k_inv_fwd_update(const SolveDesc *__restrict__ descs,
                                                        int step) {
  const SolveDesc &d = descs[blockIdx.y];
  const int nrest = d.nt - 1 - step;
  const int x = blockIdx.x;
  if (nrest <= 0 || x >= nrest * (step + 1)) return;
  const int t = step + 1 + x % nrest, c = x / nrest;
  const long ld = d.ld;
  const GZ *M = (const GZ *)d.M;
  GZ *Y = (GZ *)d.Minv;
  tile_update(M + (long)(kTile * t) + ld * (kTile * step), Y + (long)(kTile * step) + ld * (kTile * c),
              Y + (long)(kTile * t) + ld * (kTile * c), ld);
}

// Backward step s, diagonal part: Y_sc = U_ss^-1 Y_sc for every column tile c.
__global__ __launch_bounds__(256) void k_inv_bwd_panel(const SolveDesc *__restrict__ descs,
                                                       int step) {
  const SolveDesc &d = descs[blockIdx.y];
  const int c = blockIdx.x;
  if (step >= d.nt || c >= d.nt) return;
  __shared__ Z Ts[kTile][kTile + 1];
  const int tid = threadIdx.x, w = tid >> 6;
  const int wr = 32 * (w & 1), wc = 32 * (w >> 1);
  const long ld = d.ld;
  GZ *tile = (GZ *)d.Minv + (long)(step * kTile) + ld * (c * kTile);
  const GZ *__restrict__ inv = (const GZ *)d.inv + step * d.inv_stride;
  for (int e = tid; e < kTile * kTile; e += 256) {
    int r = e & 63, col = e >> 6;
    Ts[col][r] = tile[r + ld * col];
  }
  __syncthreads();
  Acc3<double> q;
  zero_acc(q);
  // A(i,k) = U^-1(wr+i, k), upper incl. diagonal; B(k,j) = Y_sc(k, wc+j)
  wave_gemm3m<double, false>([&](int i, int k) {
                               int ii = wr + i;
                               Z v = Z{0.0, 0.0};
                               if (k >= ii) v = inv[ii + kTile * k];
                               return v;
                             },
                             [&](int k, int j) { return Ts[wc + j][k]; }, q);
  store_acc(tile + wr + ld * wc, ld, q);
}

// Backward step s, update: Y_tc -= Ut_ts Y_sc for t < s, all c.
__global__ __launch_bounds__(256) void k_inv_bwd_update(const SolveDesc *__restrict__ descs,
                                                        int step) {
  const SolveDesc &d = descs[blockIdx.y];
  const int x = blockIdx.x;
  if (step >= d.nt || step == 0 || x >= step * d.nt) return;
  const int t = x % step, c = x / step;
  const long ld = d.ld;
  const GZ *M = (const GZ *)d.M;
  GZ *Y = (GZ *)d.Minv;
  tile_update(M + (long)(kTile * t) + ld * (kTile * step), Y + (long)(kTile * step) + ld * (kTile * c),
              Y + (long)(kTile * t) + ld * (kTile * c), ld);
}

// X = M^-1 [0; t0] = Y[:, central] t0 (one 64-row tile per workgroup) and
// tau00 = X_c, the central rows.
__global__ __launch_bounds__(256) void k_inv_x(const SolveDesc *__restrict__ descs) {
  const SolveDesc &d = descs[blockIdx.y];
  const int tile = blockIdx.x;
  if (tile >= d.nt) return;
  const int kc = d.kc, Ns = d.Ns;
  const long ld = d.ld;
  const Z *Y = d.Minv;
  for (int e = threadIdx.x; e < kTile * kc; e += 256) {
    const int r = tile * kTile + (e & 63), col = e >> 6;
    Z s = Z{0.0, 0.0};
    for (int k = 0; k < kc; k++) {
      const Z a = Y[r + ld * (Ns - kc + k)];
      const Z b = d.t0[k + (long)d.ldt * col];
      s.x += a.x * b.x - a.y * b.y;
      s.y += a.x * b.y + a.y * b.x;
    }
    d.X[r + (long)Ns * col] = s;
    if (r >= Ns - kc) d.tau00[(r - (Ns - kc)) + kc * col] = s;
  }
}

// Full tau in reference order: tau(rr, cr) = sum_k Minv(rs, off_j + k) t_j(k, cc),
// one solver-layout column cs = blockIdx.x per workgroup.  colsite[cs] is the
// LIZ site of the column (-1 in the identity padding), ref_of[x] the reference
// index of solver index x; t_j is the site's t packed as in the assembly.
__global__ __launch_bounds__(256) void k_tau_full_out(const SolveDesc *__restrict__ descs,
                                                      const double2 *__restrict__ tmat, long ldT,
                                                      const int *__restrict__ colsite,
                                                      const int *__restrict__ ref_of,
                                                      double2 *__restrict__ out, long ldo) {
  const SolveDesc &d = descs[0];
  const int cs = blockIdx.x;
  if (cs >= d.Ns) return;
  const int j = colsite[cs];
  if (j < 0) return;
  const int nsc = d.n_spin_cant, kkrsz = d.kkrsz;
  const int lmj = d.lmax[j], kkrj = (lmj + 1) * (lmj + 1), kkrj_ns = nsc * kkrj;
  const int offj = d.site_off[j], cc = cs - offj;
  const int csp = cc / kkrj, ci = cc % kkrj;
  const double2 *T = tmat + ldT * d.store[j] + d.t_off;
  __shared__ Z tcol[2 * kMaxKkr];
  if (threadIdx.x < kkrj_ns) {
    const int k = threadIdx.x, rsp = k / kkrj, ri = k % kkrj;
    tcol[k] = T[(rsp * kkrsz + ri) + (long)d.ldt * (csp * kkrsz + ci)];
  }
  __syncthreads();
  const long ld = d.ld;
  const long oc = ldo * ref_of[cs];
  for (int rs = d.pad + threadIdx.x; rs < d.Ns; rs += 256) {
    Z s = Z{0.0, 0.0};
    for (int k = 0; k < kkrj_ns; k++) {
      const Z a = d.Minv[rs + ld * (offj + k)];
      const Z b = tcol[k];
      s.x += a.x * b.x - a.y * b.y;
      s.y += a.x * b.y + a.y * b.x;
    }
    out[ref_of[rs] + oc] = s;
  }
}

// Sum of the per-site parts of the Lloyd trace, in site order.
__global__ void k_lloyd_reduce(const SolveDesc *__restrict__ descs, int n, double2 *out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const SolveDesc &d = descs[b];
  double2 s = make_double2(0.0, 0.0);
  for (int i = 0; i < d.numLIZ; i++) {
    s.x += d.lpart[i].x;
    s.y += d.lpart[i].y;
  }
  out[b] = s;
}

// Cotangent of tau00 -> gradient with respect to site i's t (blockIdx.x) of
// solve blockIdx.y, in the conjugate (vector-Jacobian) convention:
//   tbar_i = R_i^H V (C_i + delta_i0 I)^H,  R_i = Minv(central rows, site i),
//   C_i = gx rows of site i.
// Written unpacked into a kc x kc block with the store's layout (rows/cols
// rs*kkrsz + ri of each spin block; entries beyond a reduced lmax are 0).
__global__ __launch_bounds__(256) void k_tau_grad(const SolveDesc *__restrict__ descs,
                                                  const double2 *__restrict__ vbar,
                                                  double2 *__restrict__ gblk, int maxliz) {
  const int b = blockIdx.y, i = blockIdx.x;
  const SolveDesc &d = descs[b];
  if (i >= d.numLIZ) return;
  const int kc = d.kc, Ns = d.Ns, nsc = d.n_spin_cant, kkrsz = d.kkrsz;
  const int lmi = d.lmax[i], kkri = (lmi + 1) * (lmi + 1), kkri_ns = nsc * kkri;
  const int offi = d.site_off[i];
  const long ld = d.ld;
  __shared__ Z Rs[2 * kMaxKkr][2 * kMaxKkr + 1];  // R(a, q)
  __shared__ Z Cs[2 * kMaxKkr][2 * kMaxKkr + 1];  // C(p, col) (+ I at the central site)
  __shared__ Z Vs[2 * kMaxKkr][2 * kMaxKkr + 1];  // V(a, col)
  __shared__ Z Qs[2 * kMaxKkr][2 * kMaxKkr + 1];  // Q(a, p) = V (C + dI)^H
  const int tid = threadIdx.x;
  const double2 *V = vbar + (long)b * kc * kc;
  for (int e = tid; e < kc * kkri_ns; e += 256) {
    const int a = e % kc, q = e / kc;
    Rs[a][q] = d.Minv[(Ns - kc + a) + ld * (offi + q)];
  }
  for (int e = tid; e < kkri_ns * kc; e += 256) {
    const int p = e % kkri_ns, col = e / kkri_ns;
    Z c = d.gx[(offi + p) + (long)Ns * col];
    if (i == 0 && p == col) c.x += 1.0;
    Cs[p][col] = c;
  }
  for (int e = tid; e < kc * kc; e += 256) Vs[e % kc][e / kc] = V[e];
  __syncthreads();
  for (int e = tid; e < kc * kkri_ns; e += 256) {
    const int a = e % kc, p = e / kc;
    Z s = Z{0.0, 0.0};
    for (int col = 0; col < kc; col++) {  // V(a,col) conj(C(p,col))
      const Z v = Vs[a][col], c = Cs[p][col];
      s.x += v.x * c.x + v.y * c.y;
      s.y += v.y * c.x - v.x * c.y;
    }
    Qs[a][p] = s;
  }
  __syncthreads();
  double2 *G = gblk + ((long)b * maxliz + i) * kc * kc;
  for (int e = tid; e < kc * kc; e += 256) {
    const int r = e % kc, c = e / kc;
    const int rsp = r / kkrsz, ri = r % kkrsz, csp = c / kkrsz, ci = c % kkrsz;
    Z s = Z{0.0, 0.0};
    if (ri < kkri && ci < kkri) {
      const int q = rsp * kkri + ri, p = csp * kkri + ci;
      for (int a = 0; a < kc; a++) {  // conj(R(a,q)) Q(a,p)
        const Z rr = Rs[a][q], qq = Qs[a][p];
        s.x += rr.x * qq.x + rr.y * qq.y;
        s.y += rr.x * qq.y - rr.y * qq.x;
      }
    }
    G[e] = s;
  }
}

// Deterministic scatter-add of the per-(solve, site) gradient blocks into the
// store-shaped gradient: one workgroup per key (store column, slot offset),
// entries summed in the host-given order.
__global__ __launch_bounds__(256) void k_grad_reduce(const double2 *__restrict__ gblk, int kc,
                                                     const int *__restrict__ key_store,
                                                     const long *__restrict__ key_toff,
                                                     const int *__restrict__ key_ptr,
                                                     const int *__restrict__ entries,
                                                     double2 *__restrict__ gstore, long ldT,
                                                     int ldt) {
  const int k = blockIdx.x;
  double2 *out = gstore + ldT * key_store[k] + key_toff[k];
  for (int e = threadIdx.x; e < kc * kc; e += 256) {
    double2 s = make_double2(0.0, 0.0);
    for (int p = key_ptr[k]; p < key_ptr[k + 1]; p++) {
      const double2 v = gblk[(long)entries[p] * kc * kc + e];
      s.x += v.x;
      s.y += v.y;
    }
    double2 &o = out[(e % kc) + (long)ldt * (e / kc)];
    o.x += s.x;
    o.y += s.y;
  }
}

// Zero n complex values.
__global__ void k_zero_c(double2 *p, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    p[e] = make_double2(0.0, 0.0);
}

}  // namespace lsmsx
