// kkr_assembly.hip -- KKR matrix assembly M = 1 - t_i G_ij(E) on MI355X.
//
// Math of the reference's F77 route (calculateTauMatrix.cpp:57-315 driving
// MultipleScattering/makegij_c.f and setgij.f):
//   hfn_l  = -i^(l+1) kappa h_l(kappa r)           (makegij_c.f:139-163)
//   D_lm   = hfn_l Pbar_l|m|(cos t) e^{-+i|m|phi} (makegij_c.f:186-214)
//   G(L2,L1) = 4pi i^(l2-l1) sum_l3 cgnt(l3/2,L1,L2) D_{l3,m2-m1}
//   M(i,j) = -tpack_i * diag_spin(G_ij),  M(i,i) = 1
// Design: one workgroup per (row site i, solve) with t_i and the angular
// tables in LDS; each wave builds (i, j) blocks for its own column sites with
// wave-scope synchronisation only.  D is computed lane-parallel (one (l,m)
// per lane), and every lane writes consecutive rows of a column so the HBM
// stores (the kernel's roofline) are coalesced.
#include "lsmsx_device.hpp"

namespace lsmsx {

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cdiv(double2 a, double2 b) {
  double s = 1.0 / (b.x * b.x + b.y * b.y);
  return make_double2((a.x * b.x + a.y * b.y) * s, (a.y * b.x - a.x * b.y) * s);
}
__device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline double2 ipow_c(int n) {
  switch (n & 3) {
    case 0: return make_double2(1.0, 0.0);
    case 1: return make_double2(0.0, 1.0);
    case 2: return make_double2(-1.0, 0.0);
    default: return make_double2(0.0, -1.0);
  }
}

// hfn_l for one l (makegij_c.f:139-163), prel == 0 -> static limit.
__device__ double2 hankel_l(double2 prel, double r, int l) {
  if (prel.x == 0.0 && prel.y == 0.0) {
    double2 h = make_double2(1.0 / r, 0.0);
    for (int k = 1; k <= l; k++) h = cscale(cmul(make_double2(0.0, 1.0), h), (2.0 * k - 1.0) / r);
    return h;
  }
  double2 z = make_double2(prel.x * r, prel.y * r);
  double2 pm = make_double2(0.0, -1.0);                      // p_0 = -i
  double2 iz = cdiv(make_double2(0.0, 1.0), z);
  double2 p = make_double2(-1.0 - iz.x, -iz.y);              // p_1 = -1 - i/z
  double2 pl = (l == 0) ? pm : p;
  for (int k = 1; k < l; k++) {
    double2 t = cdiv(cscale(p, 2.0 * k + 1.0), z);
    double2 pn = make_double2(t.x - pm.x, t.y - pm.y);
    pm = p;
    p = pn;
    pl = p;
  }
  // e^{iz}/r
  double ex = exp(-z.y) / r;
  double s, c;
  sincos(z.x, &s, &c);
  double2 f = make_double2(ex * c, ex * s);
  double2 h = cmul(cmul(pl, f), ipow_c(l + 1));
  return make_double2(-h.x, -h.y);
}

// Pbar_{l,m} (m >= 0) by the recursions of Misc/associatedLegendreFunction.hpp:33-69.
__device__ double plm_bar(int l, int m, double x) {
  const double pi = 3.141592653589793238462643383279502884;
  double y = sqrt(1.0 - x * x);
  double pmm = sqrt(1.0 / (4.0 * pi));
  for (int k = 1; k <= m; k++) pmm = -sqrt((2.0 * k + 1.0) / (2.0 * k)) * y * pmm;
  if (l == m) return pmm;
  double p1 = sqrt(2.0 * m + 3.0) * x * pmm;  // Pbar_{m+1,m}
  if (l == m + 1) return p1;
  double p0 = pmm;
  for (int ll = m + 2; ll <= l; ll++) {
    double a = sqrt((4.0 * ll * ll - 1.0) / (double)(ll * ll - m * m));
    double b = sqrt(((ll - 1.0) * (ll - 1.0) - m * m) / (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
    double pn = a * (x * p1 - b * p0);
    p0 = p1;
    p1 = pn;
  }
  return p1;
}

// Wave-scope ordering of LDS traffic: every lane's earlier LDS writes are
// visible to the wave's later LDS reads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int kLmaxD = 2 * kMaxLmax;                       // l <= lend of D
constexpr int kNd = (kLmaxD + 1) * (kLmaxD + 1);           // 49 (l,m) of D

// Assemble row-site block row i = blockIdx.x of solve blockIdx.y:
// M(i,j) = -tpack_i diag_spin(G_ij) for j != i, identity for j == i.
// The workgroup stages the packed t_i and the angular tables in LDS; each
// wave then builds blocks (i, j), j = w, w+4, ..., independently.
__global__ __launch_bounds__(256) void k_assemble(const SolveDesc *__restrict__ descs,
                                                  const double2 *__restrict__ tmat, long ldT,
                                                  const double *__restrict__ cgnt_g,
                                                  const double2 *__restrict__ illp_g,
                                                  int lmax_tab) {
  const SolveDesc &d = descs[blockIdx.y];
  const int i = blockIdx.x;
  if (i >= d.numLIZ) return;

  __shared__ double cg[(kMaxLmax + 1) * kMaxKkr * kMaxKkr];
  __shared__ double2 il[kMaxKkr * kMaxKkr];
  __shared__ double2 tn[2 * kMaxKkr][2 * kMaxKkr + 1];   // t_i in 16-aligned spin blocks, [row][col]
  __shared__ double pa[kLmaxD + 1][kLmaxD + 1], pb[kLmaxD + 1][kLmaxD + 1];
  __shared__ double pmmf[kLmaxD + 1], pm1f[kLmaxD + 1];
  __shared__ int lofk[kNd];
  __shared__ double2 dl[4][kNd];
  __shared__ double2 gs[4][kMaxKkr * kMaxKkr];

  const int tid = threadIdx.x;
  const int kkr_tab = (lmax_tab + 1) * (lmax_tab + 1);
  const int nsc = d.n_spin_cant, kkrsz = d.kkrsz, ldt = d.ldt;
  const int lmi = d.lmax[i], kkri = (lmi + 1) * (lmi + 1), kkri_ns = kkri * nsc;
  for (int t = tid; t < (lmax_tab + 1) * kkr_tab * kkr_tab; t += 256) cg[t] = cgnt_g[t];
  for (int t = tid; t < kkr_tab * kkr_tab; t += 256) il[t] = illp_g[t];
  // t_i packed as the F77 route does (calculateTauMatrix.cpp:155-183: the
  // leading kkri block of each spin block), each spin block placed at a
  // 16-aligned offset and zero-padded for the MFMA product below.
  {
    const double2 *__restrict__ T = tmat + ldT * d.store[i] + d.t_off;
    for (int t = tid; t < 4 * kMaxKkr * kMaxKkr; t += 256) {
      int r = t % (2 * kMaxKkr), c = t / (2 * kMaxKkr);
      int rs = r / kMaxKkr, ri = r % kMaxKkr, cs = c / kMaxKkr, ci = c % kMaxKkr;
      double2 v = make_double2(0.0, 0.0);
      if (rs < nsc && cs < nsc && ri < kkri && ci < kkri)
        v = T[(rs * kkrsz + ri) + (long)ldt * (cs * kkrsz + ci)];
      tn[r][c] = v;
    }
  }
  // Pbar recursion coefficients (Misc/associatedLegendreFunction.hpp:33-69)
  if (tid < (kLmaxD + 1) * (kLmaxD + 1)) {
    int l = tid / (kLmaxD + 1), m = tid % (kLmaxD + 1);
    double a = 0.0, b = 0.0;
    if (l >= m + 2) {
      a = sqrt((4.0 * l * l - 1.0) / (double)(l * l - m * m));
      b = sqrt(((l - 1.0) * (l - 1.0) - m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
    }
    pa[l][m] = a;
    pb[l][m] = b;
  }
  if (tid < kLmaxD + 1) {
    pmmf[tid] = tid >= 1 ? sqrt((2.0 * tid + 1.0) / (2.0 * tid)) : 1.0;
    pm1f[tid] = sqrt(2.0 * tid + 3.0);
  }
  if (tid < kNd) {
    int l = 0;
    while ((l + 1) * (l + 1) <= tid) l++;
    lofk[tid] = l;
  }
  __syncthreads();

  const int w = tid >> 6, lane = tid & 63;
  const double xi = d.pos[3 * i], yi = d.pos[3 * i + 1], zi = d.pos[3 * i + 2];
  const double2 prel = make_double2(d.prel_re, d.prel_im);
  const bool prel0 = (prel.x == 0.0 && prel.y == 0.0);
  const double pi4 = 4.0 * 3.141592653589793238462643383279502884;
  const double y00 = 0.28209479177387814347;  // 1/sqrt(4 pi)
  double2 *__restrict__ M = d.M;
  // mixed precision: the solver factors an FP32 copy written alongside M
  float2 *__restrict__ W = (d.A != (void *)d.M) ? (float2 *)d.A : nullptr;
  const long ld = d.ld;
  const int offi = d.site_off[i];

  for (int j = w; j < d.numLIZ; j += 4) {
    const int lmj = d.lmax[j], kkrj = (lmj + 1) * (lmj + 1), kkrj_ns = kkrj * nsc;
    const int offj = d.site_off[j];
    if (j == i) {
      for (int e = lane; e < kkri_ns * kkri_ns; e += 64) {
        int rr = e % kkri_ns, cc = e / kkri_ns;
        M[(offi + rr) + ld * (offj + cc)] = make_double2(rr == cc ? 1.0 : 0.0, 0.0);
        if (W) W[(offi + rr) + ld * (offj + cc)] = make_float2(rr == cc ? 1.f : 0.f, 0.f);
      }
      continue;
    }
    const double rx = xi - d.pos[3 * j], ry = yi - d.pos[3 * j + 1], rz = zi - d.pos[3 * j + 2];
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    const int lend = lmi + lmj;
    // --- D_{l,m}, one (l,m) per lane (makegij_c.f:139-214)
    if (lane < (lend + 1) * (lend + 1)) {
      const int l = lofk[lane], m = lane - l * (l + 1), am = m < 0 ? -m : m;
      double2 h;
      if (prel0) {  // static limit: hfn_l = i^l (2l-1)!! / r^(l+1)
        double v = 1.0 / r;
        for (int k = 1; k <= l; k++) v *= (2.0 * k - 1.0) / r;
        h = cmul(ipow_c(l), make_double2(v, 0.0));
      } else {
        // p_0 = -i, p_1 = -1 - i/z, p_{k+1} = (2k+1) p_k / z - p_{k-1};
        // hfn_l = -p_l e^{iz}/r i^{l+1}
        const double2 z = make_double2(prel.x * r, prel.y * r);
        const double zn = 1.0 / (z.x * z.x + z.y * z.y);
        const double2 iz = make_double2(z.x * zn, -z.y * zn);  // 1/z
        double2 pm = make_double2(0.0, -1.0);
        double2 p = make_double2(-1.0 + iz.y, -iz.x);          // -1 - i/z
        if (l == 0) p = pm;
        for (int k = 1; k < l; k++) {
          double2 t = cscale(cmul(p, iz), 2.0 * k + 1.0);
          double2 pn = make_double2(t.x - pm.x, t.y - pm.y);
          pm = p;
          p = pn;
        }
        const double ex = exp(-z.y) / r;
        double sn, cs;
        sincos(z.x, &sn, &cs);
        h = cmul(cmul(p, make_double2(ex * cs, ex * sn)), ipow_c(l + 1));
        h = make_double2(-h.x, -h.y);
      }
      // Pbar_{l,|m|}(cos theta)
      const double x = rz / r, y = sqrt(1.0 - x * x);
      double pmm = y00;
      for (int k = 1; k <= am; k++) pmm = -pmmf[k] * y * pmm;
      double pv = pmm;
      if (l > am) {
        double p0 = pmm, p1 = pm1f[am] * x * pmm;
        for (int ll = am + 2; ll <= l; ll++) {
          double pn = pa[ll][am] * (x * p1 - pb[ll][am] * p0);
          p0 = p1;
          p1 = pn;
        }
        pv = p1;
      }
      // cos/sin(|m| phi), zero azimuth on the z axis
      const double pmag = sqrt(rx * rx + ry * ry);
      double c1 = 0.0, s1 = 0.0;
      if (pmag >= 1.0e-6) {
        c1 = rx / pmag;
        s1 = ry / pmag;
      }
      double cm = 1.0, sm = 0.0;
      for (int k = 1; k <= am; k++) {
        double cn = cm * c1 - sm * s1;
        double snn = sm * c1 + cm * s1;
        cm = cn;
        sm = snn;
      }
      const double2 fac = make_double2(pv * cm, pv * sm);
      double2 v;
      if (m == 0) v = cscale(h, pv);
      else if (m < 0) v = cscale(cmul(h, fac), (am & 1) ? -1.0 : 1.0);
      else v = cmul(h, make_double2(fac.x, -fac.y));
      dl[w][lane] = v;
    }
    wave_sync();
    // --- G_ij(L2, L1) (makegij_c.f:240-263), stored 16 x 16 zero-padded
    for (int e = lane; e < kMaxKkr * kMaxKkr; e += 64) {
      const int lm2 = e % kMaxKkr, lm1 = e / kMaxKkr;
      double2 gv = make_double2(0.0, 0.0);
      if (lm2 < kkri && lm1 < kkrj) {
        const int l1 = lofk[lm1], l2 = lofk[lm2];
        const int m1 = lm1 - l1 * (l1 + 1), m2 = lm2 - l2 * (l2 + 1), m3 = m2 - m1;
        const int am3 = m3 < 0 ? -m3 : m3, dl12 = l1 - l2 < 0 ? l2 - l1 : l1 - l2;
        int llow = am3 > dl12 ? am3 : dl12;
        if (prel0) llow = l1 + l2;
        double2 sacc = make_double2(0.0, 0.0);
        for (int l3 = l1 + l2; l3 >= llow; l3 -= 2) {
          const double c = cg[l3 / 2 + (lmax_tab + 1) * (lm1 + kkr_tab * lm2)];
          const double2 dv = dl[w][l3 * (l3 + 1) + m3];
          sacc.x += c * dv.x;
          sacc.y += c * dv.y;
        }
        gv = cscale(cmul(il[lm2 + kkr_tab * lm1], sacc), pi4);
      }
      gs[w][lm2 + kMaxKkr * lm1] = gv;
    }
    wave_sync();
    // --- M block = -tpack_i diag_spin(G) (calculateTauMatrix.cpp:240-242) on
    // FP64 MFMA, 3M form, one 16 x 16 x 16 product per spin block pair
    // (bi, bj): M_bi,bj = -t_bi,bj G.  Transposed product as in the LU update,
    // so register r of lane (li, kq) is element (li, kq + 4r) of the block and
    // each store is a run of 16 consecutive rows.
    {
      const int li = lane & 15, kq = lane >> 4;
      for (int bi = 0; bi < nsc; bi++)
        for (int bj = 0; bj < nsc; bj++) {
          d4 q1 = {0.0, 0.0, 0.0, 0.0}, q2 = q1, q3 = q1;
#pragma unroll
          for (int k0 = 0; k0 < kMaxKkr; k0 += 4) {
            const double2 g = gs[w][(k0 + kq) + kMaxKkr * li];
            const double2 t = tn[bi * kMaxKkr + li][bj * kMaxKkr + k0 + kq];
            q1 = __builtin_amdgcn_mfma_f64_16x16x4f64(g.x, t.x, q1, 0, 0, 0);
            q2 = __builtin_amdgcn_mfma_f64_16x16x4f64(g.y, t.y, q2, 0, 0, 0);
            q3 = __builtin_amdgcn_mfma_f64_16x16x4f64(g.x + g.y, t.x + t.y, q3, 0, 0, 0);
          }
          if (li < kkri) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int jc = kq + 4 * r;
              if (jc < kkrj) {
                const double re = q1[r] - q2[r], im = q3[r] - q1[r] - q2[r];
                const long idx = (offi + bi * kkri + li) + ld * (offj + bj * kkrj + jc);
                M[idx] = make_double2(-re, -im);
                if (W) W[idx] = make_float2((float)-re, (float)-im);
              }
            }
          }
        }
    }
    wave_sync();
  }
}

// Identity in the leading pad rows/columns of the solver layout.
__global__ __launch_bounds__(256) void k_pad_identity(const SolveDesc *__restrict__ descs) {
  const SolveDesc &d = descs[blockIdx.x];
  const int pad = d.pad;
  if (pad == 0) return;
  double2 *M = d.M;
  float2 *W = (d.A != (void *)d.M) ? (float2 *)d.A : nullptr;
  const long ld = d.ld;
  // columns 0..pad-1: full columns; columns pad..Ns-1: rows 0..pad-1
  for (long e = threadIdx.x; e < (long)pad * d.Ns; e += blockDim.x) {
    int rr = (int)(e % d.Ns), cc = (int)(e / d.Ns);
    M[rr + ld * cc] = make_double2(rr == cc ? 1.0 : 0.0, 0.0);
    if (W) W[rr + ld * cc] = make_float2(rr == cc ? 1.f : 0.f, 0.f);
  }
  for (long e = threadIdx.x; e < (long)pad * (d.Ns - pad); e += blockDim.x) {
    int rr = (int)(e % pad), cc = pad + (int)(e / pad);
    M[rr + ld * cc] = make_double2(0.0, 0.0);
    if (W) W[rr + ld * cc] = make_float2(0.f, 0.f);
  }
}

}  // namespace lsmsx
