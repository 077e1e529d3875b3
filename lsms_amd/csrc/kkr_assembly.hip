// kkr_assembly.hip -- KKR matrix assembly M = 1 - t_i G_ij(E) on MI355X.
//
// Math of the reference's F77 route (calculateTauMatrix.cpp:57-315 driving
// MultipleScattering/makegij_c.f and setgij.f):
//   hfn_l  = -i^(l+1) kappa h_l(kappa r)           (makegij_c.f:139-163)
//   D_lm   = hfn_l Pbar_l|m|(cos t) e^{-+i|m|phi} (makegij_c.f:186-214)
//   G(L2,L1) = 4pi i^(l2-l1) sum_l3 cgnt(l3/2,L1,L2) D_{l3,m2-m1}
//   M(i,j) = -tpack_i * diag_spin(G_ij),  M(i,i) = 1
// Design: one workgroup per (column site j, solve); each wave builds one
// (i, j) block at a time.  D is computed lane-parallel (one (l,m) per lane),
// the Gaunt/i^l tables sit in LDS, and every lane writes consecutive rows of
// a column so the HBM stores (the kernel's roofline) are coalesced.
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

// Assemble all blocks of column site j = blockIdx.x for solve blockIdx.y.
// Off-diagonal blocks: -tpack_i diag_spin(G_ij); diagonal blocks: identity.
__global__ __launch_bounds__(256) void k_assemble(const SolveDesc *__restrict__ descs,
                                                  const double2 *__restrict__ tmat, long ldT,
                                                  const double *__restrict__ cgnt_g,
                                                  const double2 *__restrict__ illp_g,
                                                  int lmax_tab) {
  const SolveDesc &d = descs[blockIdx.y];
  const int j = blockIdx.x;
  if (j >= d.numLIZ) return;

  __shared__ double cg[(kMaxLmax + 1) * kMaxKkr * kMaxKkr];
  __shared__ double2 il[kMaxKkr * kMaxKkr];
  __shared__ double2 dl[4][(2 * kMaxLmax + 1) * (2 * kMaxLmax + 1)];
  __shared__ double2 gs[4][kMaxKkr * kMaxKkr];

  const int kkr_tab = (lmax_tab + 1) * (lmax_tab + 1);
  for (int t = threadIdx.x; t < (lmax_tab + 1) * kkr_tab * kkr_tab; t += blockDim.x) cg[t] = cgnt_g[t];
  for (int t = threadIdx.x; t < kkr_tab * kkr_tab; t += blockDim.x) il[t] = illp_g[t];
  __syncthreads();

  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nsc = d.n_spin_cant;
  const int lmj = d.lmax[j], kkrj = (lmj + 1) * (lmj + 1), kkrj_ns = kkrj * nsc;
  const int offj = d.site_off[j];
  const double xj = d.pos[3 * j], yj = d.pos[3 * j + 1], zj = d.pos[3 * j + 2];
  const double2 prel = make_double2(d.prel_re, d.prel_im);
  const bool prel0 = (prel.x == 0.0 && prel.y == 0.0);
  const double pi4 = 4.0 * 3.141592653589793238462643383279502884;
  double2 *__restrict__ M = d.M;
  // mixed precision: the solver factors an FP32 copy written alongside M
  float2 *__restrict__ W = (d.A != (void *)d.M) ? (float2 *)d.A : nullptr;
  const long ld = d.ld;

  for (int i0 = 0; i0 < d.numLIZ; i0 += 4) {
    const int i = i0 + w;
    const bool act = i < d.numLIZ;
    int lmi = 0, kkri = 1, kkri_ns = 1, offi = 0, lend = 0;
    double rx = 0, ry = 0, rz = 0, r = 1;
    if (act) {
      lmi = d.lmax[i];
      kkri = (lmi + 1) * (lmi + 1);
      kkri_ns = kkri * nsc;
      offi = d.site_off[i];
      rx = d.pos[3 * i] - xj;
      ry = d.pos[3 * i + 1] - yj;
      rz = d.pos[3 * i + 2] - zj;
      r = sqrt(rx * rx + ry * ry + rz * rz);
      lend = lmi + lmj;
    }
    const bool offdiag = act && (i != j);
    // --- D_{l,m}, one (l,m) per lane (makegij_c.f:186-214)
    if (offdiag && lane < (lend + 1) * (lend + 1)) {
      int l = (int)sqrt((double)lane + 0.5);
      while (l * l > lane) l--;
      while ((l + 1) * (l + 1) <= lane) l++;
      int m = lane - l * (l + 1), am = m < 0 ? -m : m;
      double2 h = hankel_l(prel, r, l);
      double p = plm_bar(l, am, rz / r);
      double pmag = sqrt(rx * rx + ry * ry);
      double c1 = 0.0, s1 = 0.0;
      if (pmag >= 1.0e-6) {
        c1 = rx / pmag;
        s1 = ry / pmag;
      }
      double cm = 1.0, sm = 0.0;
      for (int k = 1; k <= am; k++) {
        double cn = cm * c1 - sm * s1;
        double sn = sm * c1 + cm * s1;
        cm = cn;
        sm = sn;
      }
      double2 fac = make_double2(p * cm, p * sm);
      double2 v;
      if (m == 0) v = cscale(h, p);
      else if (m < 0) v = cscale(cmul(h, fac), (am & 1) ? -1.0 : 1.0);
      else v = cmul(h, make_double2(fac.x, -fac.y));
      dl[w][lane] = v;
    }
    __syncthreads();
    // --- G_ij(L2, L1) (makegij_c.f:240-263)
    if (offdiag) {
      for (int e = lane; e < kkri * kkrj; e += 64) {
        int lm2 = e % kkri, lm1 = e / kkri;
        int l1 = (int)sqrt((double)lm1 + 0.5);
        while (l1 * l1 > lm1) l1--;
        while ((l1 + 1) * (l1 + 1) <= lm1) l1++;
        int l2 = (int)sqrt((double)lm2 + 0.5);
        while (l2 * l2 > lm2) l2--;
        while ((l2 + 1) * (l2 + 1) <= lm2) l2++;
        int m1 = lm1 - l1 * (l1 + 1), m2 = lm2 - l2 * (l2 + 1), m3 = m2 - m1;
        int am3 = m3 < 0 ? -m3 : m3, dl12 = l1 - l2 < 0 ? l2 - l1 : l1 - l2;
        int llow = am3 > dl12 ? am3 : dl12;
        if (prel0) llow = l1 + l2;
        double2 s = make_double2(0.0, 0.0);
        for (int l3 = l1 + l2; l3 >= llow; l3 -= 2) {
          double c = cg[l3 / 2 + (lmax_tab + 1) * (lm1 + kkr_tab * lm2)];
          double2 dv = dl[w][l3 * (l3 + 1) + m3];
          s.x += c * dv.x;
          s.y += c * dv.y;
        }
        gs[w][lm2 + kkri * lm1] = cscale(cmul(il[lm2 + kkr_tab * lm1], s), pi4);
      }
    }
    __syncthreads();
    // --- M block = -tpack_i * diag_spin(G) (calculateTauMatrix.cpp:155-183,240-242)
    if (offdiag) {
      const double2 *__restrict__ T = tmat + ldT * d.store[i] + d.t_off;
      const int ldt = d.ldt, kkrsz = d.kkrsz;
      for (int e = lane; e < kkri_ns * kkrj_ns; e += 64) {
        int rr = e % kkri_ns, cc = e / kkri_ns;
        int isr = rr / kkri, ii = rr % kkri;
        int cs = cc / kkrj, jj = cc % kkrj;
        const double2 *Trow = T + (isr * kkrsz + ii) + (long)ldt * (cs * kkrsz);
        double2 acc = make_double2(0.0, 0.0);
        for (int k = 0; k < kkri; k++) {
          double2 tv = Trow[(long)ldt * k];
          double2 gv = gs[w][k + kkri * jj];
          acc.x += tv.x * gv.x - tv.y * gv.y;
          acc.y += tv.x * gv.y + tv.y * gv.x;
        }
        M[(offi + rr) + ld * (offj + cc)] = make_double2(-acc.x, -acc.y);
        if (W) W[(offi + rr) + ld * (offj + cc)] = make_float2((float)-acc.x, (float)-acc.y);
      }
    } else if (act) {
      for (int e = lane; e < kkri_ns * kkri_ns; e += 64) {
        int rr = e % kkri_ns, cc = e / kkri_ns;
        M[(offi + rr) + ld * (offj + cc)] = make_double2(rr == cc ? 1.0 : 0.0, 0.0);
        if (W) W[(offi + rr) + ld * (offj + cc)] = make_float2(rr == cc ? 1.f : 0.f, 0.f);
      }
    }
    __syncthreads();
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
