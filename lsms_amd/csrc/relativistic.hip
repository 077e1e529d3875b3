// relativistic.hip -- the fully relativistic (kappa-mu) form of the KKR
// blocks, relativity = full (setgij.f:39-55, relmtrx.f).
//
// The reference builds every off-diagonal block in the kappa-mu basis,
//   M_ij = -t_i (prel^2/E) relmtrx(G_ij),   relmtrx(G) = U diag_spin(G) U^T,
// with U the real orthogonal Clebsch-Gordan change of basis (l m s) ->
// (kappa mu) of clebsch.f.  Since U is the same for every site, the whole
// matrix is an orthogonal similarity of a spin-canted (l m s) KKR matrix:
//   M = U~ M' U~^T,  M' = 1 - t' diag_spin(G),  t'_i = (prel^2/E) U^T t_i U,
// (U~ = diag(U, ..., U)), and the central block of the inverse follows,
//   tau00 = [M^-1 [t0;0]]_00 = (E/prel^2) U [M'^-1 [t0';0]]_00 U^T.
// So the relativistic path runs the n_spin_cant = 2 kernels unchanged on the
// transformed store t' (k_rel_tstore), and these kernels map the results
// back: tau00 (k_rel_tau) and, for dumps in the reference's basis, the
// blocks of M itself (k_rel_blocks).  U is passed as a dense kk2 x kk2
// column-major matrix (kk2 = 2 (lmax+1)^2 <= 32).
#include "lsmsx_device.hpp"

namespace lsmsx {

constexpr int kRelMax = 2 * kMaxKkr;  // 32

// U (column-major, leading dimension kk2) into LDS: u[r][c] = U(r, c).
__device__ __forceinline__ void rel_load_u(const double *__restrict__ U, int kk2,
                                           double (*u)[kRelMax + 1]) {
  for (int e = threadIdx.x; e < kk2 * kk2; e += blockDim.x) u[e % kk2][e / kk2] = U[e];
}

// t' = f U^T t U, in place, for store column blockIdx.x and energy slot
// blockIdx.y (f = prel^2 / E of that slot).
__global__ __launch_bounds__(256) void k_rel_tstore(double2 *__restrict__ tmat, long ldT, long blk,
                                                    int kk2, const double *__restrict__ U,
                                                    const double2 *__restrict__ fac) {
  __shared__ double u[kRelMax][kRelMax + 1];
  __shared__ double2 t[kRelMax][kRelMax + 1], w[kRelMax][kRelMax + 1];
  double2 *T = tmat + ldT * blockIdx.x + blk * blockIdx.y;
  const double2 f = fac[blockIdx.y];
  rel_load_u(U, kk2, u);
  for (int e = threadIdx.x; e < kk2 * kk2; e += blockDim.x) t[e % kk2][e / kk2] = T[e];
  __syncthreads();
  for (int e = threadIdx.x; e < kk2 * kk2; e += blockDim.x) {  // w = t U
    const int i = e % kk2, b = e / kk2;
    double2 s = make_double2(0.0, 0.0);
    for (int j = 0; j < kk2; j++) {
      s.x += t[i][j].x * u[j][b];
      s.y += t[i][j].y * u[j][b];
    }
    w[i][b] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kk2 * kk2; e += blockDim.x) {  // t' = f U^T w
    const int a = e % kk2, b = e / kk2;
    double2 s = make_double2(0.0, 0.0);
    for (int i = 0; i < kk2; i++) {
      s.x += u[i][a] * w[i][b].x;
      s.y += u[i][a] * w[i][b].y;
    }
    T[e] = make_double2(f.x * s.x - f.y * s.y, f.x * s.y + f.y * s.x);
  }
}

// tau <- g U tau U^T for block blockIdx.x (kc x kc, column-major), g = E / prel^2
// of the solve.
__global__ __launch_bounds__(256) void k_rel_tau(double2 *__restrict__ tau, int kc,
                                                 const double *__restrict__ U,
                                                 const double2 *__restrict__ finv) {
  __shared__ double u[kRelMax][kRelMax + 1];
  __shared__ double2 t[kRelMax][kRelMax + 1], w[kRelMax][kRelMax + 1];
  double2 *T = tau + (long)kc * kc * blockIdx.x;
  const double2 g = finv[blockIdx.x];
  rel_load_u(U, kc, u);
  for (int e = threadIdx.x; e < kc * kc; e += blockDim.x) t[e % kc][e / kc] = T[e];
  __syncthreads();
  for (int e = threadIdx.x; e < kc * kc; e += blockDim.x) {  // w = tau U^T
    const int i = e % kc, b = e / kc;
    double2 s = make_double2(0.0, 0.0);
    for (int j = 0; j < kc; j++) {
      s.x += t[i][j].x * u[b][j];
      s.y += t[i][j].y * u[b][j];
    }
    w[i][b] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kc * kc; e += blockDim.x) {  // g U w
    const int a = e % kc, b = e / kc;
    double2 s = make_double2(0.0, 0.0);
    for (int i = 0; i < kc; i++) {
      s.x += u[a][i] * w[i][b].x;
      s.y += u[a][i] * w[i][b].y;
    }
    T[e] = make_double2(g.x * s.x - g.y * s.y, g.x * s.y + g.y * s.x);
  }
}

// Block (blockIdx.x, blockIdx.y) of M (kk2 x kk2 site blocks, leading
// dimension ld) <- U block U^T: the kappa-mu form of the assembled matrix.
__global__ __launch_bounds__(256) void k_rel_blocks(double2 *__restrict__ M, long ld, int kk2,
                                                    const double *__restrict__ U) {
  __shared__ double u[kRelMax][kRelMax + 1];
  __shared__ double2 t[kRelMax][kRelMax + 1], w[kRelMax][kRelMax + 1];
  double2 *B = M + (long)kk2 * blockIdx.x + ld * (long)kk2 * blockIdx.y;
  rel_load_u(U, kk2, u);
  for (int e = threadIdx.x; e < kk2 * kk2; e += blockDim.x) t[e % kk2][e / kk2] = B[(e % kk2) + ld * (e / kk2)];
  __syncthreads();
  for (int e = threadIdx.x; e < kk2 * kk2; e += blockDim.x) {  // w = B U^T
    const int i = e % kk2, b = e / kk2;
    double2 s = make_double2(0.0, 0.0);
    for (int j = 0; j < kk2; j++) {
      s.x += t[i][j].x * u[b][j];
      s.y += t[i][j].y * u[b][j];
    }
    w[i][b] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kk2 * kk2; e += blockDim.x) {  // U w
    const int a = e % kk2, b = e / kk2;
    double2 s = make_double2(0.0, 0.0);
    for (int i = 0; i < kk2; i++) {
      s.x += u[a][i] * w[i][b].x;
      s.y += u[a][i] * w[i][b].y;
    }
    B[a + ld * b] = s;
  }
}

}  // namespace lsmsx
