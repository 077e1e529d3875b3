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

// Wave-scope ordering of LDS traffic: every lane's earlier LDS writes are
// visible to the wave's later LDS reads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef double d4 __attribute__((ext_vector_type(4)));
// site pairs of one assembly workgroup whose D values share the LDS at once
constexpr int kAsmPairs = 40;
constexpr int kLmaxD = 2 * kMaxLmax;                       // l <= lend of D
constexpr int kNd = (kLmaxD + 1) * (kLmaxD + 1);           // 49 (l,m) of D

// hfn_l = -i^(l+1) p_l(z) e^{iz}/r, z = prel r (makegij_c.f:139-163), with
// p_0 = -i, p_1 = -1 - i/z, p_{k+1} = (2k+1) p_k / z - p_{k-1}.  With dh
// given, also d hfn_l / d prel = -i^(l+1) (p_l' + i p_l) e^{iz}, p' from the
// differentiated recursion p_0' = 0, p_1' = i/z^2,
// p_{k+1}' = (2k+1)(p_k'/z - p_k/z^2) - p_{k-1}'.
__device__ __forceinline__ double2 hankel_lane(double2 prel, double r, int l, double2 *dh) {
  const double2 z = make_double2(prel.x * r, prel.y * r);
  const double zn = 1.0 / (z.x * z.x + z.y * z.y);
  const double2 iz = make_double2(z.x * zn, -z.y * zn);  // 1/z
  const double2 iz2 = cmul(iz, iz);
  double2 pm = make_double2(0.0, -1.0);
  double2 p = make_double2(-1.0 + iz.y, -iz.x);          // -1 - i/z
  double2 dpm = make_double2(0.0, 0.0);
  double2 dp = make_double2(-iz2.y, iz2.x);              // i/z^2
  if (l == 0) {
    p = pm;
    dp = dpm;
  }
  for (int k = 1; k < l; k++) {
    const double2 t = cscale(cmul(p, iz), 2.0 * k + 1.0);
    const double2 pn = make_double2(t.x - pm.x, t.y - pm.y);
    if (dh) {
      const double2 a = cmul(dp, iz), b = cmul(p, iz2);
      const double2 dpn = make_double2((2.0 * k + 1.0) * (a.x - b.x) - dpm.x,
                                       (2.0 * k + 1.0) * (a.y - b.y) - dpm.y);
      dpm = dp;
      dp = dpn;
    }
    pm = p;
    p = pn;
  }
  const double ex = exp(-z.y);
  double sn, cs;
  sincos(z.x, &sn, &cs);
  const double2 e = make_double2(ex * cs, ex * sn);      // e^{iz}
  const double2 il1 = ipow_c(l + 1);
  const double2 h = cmul(cmul(p, cscale(e, 1.0 / r)), il1);
  if (dh) {
    const double2 q = make_double2(dp.x - p.y, dp.y + p.x);  // p' + i p
    const double2 g = cmul(cmul(q, e), il1);
    *dh = make_double2(-g.x, -g.y);
  }
  return make_double2(-h.x, -h.y);
}

// Angular part of D_{l,m} (makegij_c.f:170-214): Pbar_{l,|m|}(cos theta) by
// the recursions of Misc/associatedLegendreFunction.hpp:33-69 (coefficient
// tables in LDS), cos/sin(|m| phi) with zero azimuth on the z axis, and the
// sign rule D_{l,-m} = (-1)^m h Pbar e^{+i|m|phi}, D_{l,m} = h Pbar e^{-i m phi}.
struct AngTab {
  const double (*pa)[kLmaxD + 1];
  const double (*pb)[kLmaxD + 1];
  const double *pmmf, *pm1f;
};
__device__ __forceinline__ double2 angular_lane(const AngTab &tb, int l, int m, double rx,
                                                double ry, double rz, double r) {
  const double y00 = 0.28209479177387814347;  // 1/sqrt(4 pi)
  const int am = m < 0 ? -m : m;
  const double x = rz / r, y = sqrt(1.0 - x * x);
  double pmm = y00;
  for (int k = 1; k <= am; k++) pmm = -tb.pmmf[k] * y * pmm;
  double pv = pmm;
  if (l > am) {
    double p0 = pmm, p1 = tb.pm1f[am] * x * pmm;
    for (int ll = am + 2; ll <= l; ll++) {
      double pn = tb.pa[ll][am] * (x * p1 - tb.pb[ll][am] * p0);
      p0 = p1;
      p1 = pn;
    }
    pv = p1;
  }
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
  if (m == 0) return make_double2(pv, 0.0);
  if (m < 0) return cscale(make_double2(pv * cm, pv * sm), (am & 1) ? -1.0 : 1.0);
  return make_double2(pv * cm, -pv * sm);
}

// D_{l,m}(R), l <= 6, of one site pair for the |m| values of one quarter of
// the m range (PART 0: |m| = 0, 4; 1: 1, 5; 2: 2, 6; 3: 3), written to
// dlp[l(l+1) + m] and dlp[l(l+1) - m].  The same recurrences as hankel_lane and
// angular_lane in the same order of operations, evaluated once per pair and
// |m| instead of once per (l, m): the Hankel recursion runs over all l in one
// pass and each Pbar column over its l in one pass.
template <int PART>
__device__ __forceinline__ void d_pair_part(const AngTab &tb, double2 prel, bool prel0, double rx,
                                            double ry, double rz, double r,
                                            double2 *__restrict__ dlp) {
  double2 h[kLmaxD + 1];
  if (prel0) {  // static limit: hfn_l = i^l (2l-1)!! / r^(l+1)
    double v = 1.0 / r;
    h[0] = cmul(ipow_c(0), make_double2(v, 0.0));
#pragma unroll
    for (int l = 1; l <= kLmaxD; l++) {
      v *= (2.0 * l - 1.0) / r;
      h[l] = cmul(ipow_c(l), make_double2(v, 0.0));
    }
  } else {
    const double2 z = make_double2(prel.x * r, prel.y * r);
    const double zn = 1.0 / (z.x * z.x + z.y * z.y);
    const double2 iz = make_double2(z.x * zn, -z.y * zn);  // 1/z
    const double ex = exp(-z.y);
    double sn, cs;
    sincos(z.x, &sn, &cs);
    const double2 e = make_double2(ex * cs, ex * sn);      // e^{iz}
    const double2 er = cscale(e, 1.0 / r);
    double2 pm = make_double2(0.0, -1.0);
    double2 p = make_double2(-1.0 + iz.y, -iz.x);          // -1 - i/z
    {
      const double2 t = cmul(cmul(pm, er), ipow_c(1));
      h[0] = make_double2(-t.x, -t.y);
      const double2 t1 = cmul(cmul(p, er), ipow_c(2));
      h[1] = make_double2(-t1.x, -t1.y);
    }
#pragma unroll
    for (int k = 1; k < kLmaxD; k++) {
      const double2 t = cscale(cmul(p, iz), 2.0 * k + 1.0);
      const double2 pn = make_double2(t.x - pm.x, t.y - pm.y);
      pm = p;
      p = pn;
      const double2 hh = cmul(cmul(p, er), ipow_c(k + 2));
      h[k + 1] = make_double2(-hh.x, -hh.y);
    }
  }
  const double y00 = 0.28209479177387814347;  // 1/sqrt(4 pi)
  const double x = rz / r, y = sqrt(1.0 - x * x);
  const double pmag = sqrt(rx * rx + ry * ry);
  double c1 = 0.0, s1 = 0.0;
  if (pmag >= 1.0e-6) {
    c1 = rx / pmag;
    s1 = ry / pmag;
  }
  // Pbar_{|m|,|m|} and cos/sin(|m| phi) by the recursions of angular_lane,
  // carried from |m| = part to part + 4
  double pmm = y00, cm = 1.0, sm = 0.0;
#pragma unroll
  for (int k = 1; k <= PART; k++) {
    pmm = -tb.pmmf[k] * y * pmm;
    const double cn = cm * c1 - sm * s1;
    const double snn = sm * c1 + cm * s1;
    cm = cn;
    sm = snn;
  }
#pragma unroll
  for (int am = PART; am <= kLmaxD; am += 4) {
    if (am > PART)
#pragma unroll
      for (int k = am - 3; k <= am; k++) {
        pmm = -tb.pmmf[k] * y * pmm;
        const double cn = cm * c1 - sm * s1;
        const double snn = sm * c1 + cm * s1;
        cm = cn;
        sm = snn;
      }
    const double sgn = (am & 1) ? -1.0 : 1.0;
    double p0 = pmm, p1 = 0.0;
#pragma unroll
    for (int l = am; l <= kLmaxD; l++) {
      double pv;
      if (l == am) {
        pv = pmm;
      } else if (l == am + 1) {
        p1 = tb.pm1f[am] * x * pmm;
        pv = p1;
      } else {
        const double pn = tb.pa[l][am] * (x * p1 - tb.pb[l][am] * p0);
        p0 = p1;
        p1 = pn;
        pv = p1;
      }
      const int c = l * (l + 1);
      if (am == 0) {
        dlp[c] = cmul(h[l], make_double2(pv, 0.0));
      } else {
        dlp[c + am] = cmul(h[l], make_double2(pv * cm, -pv * sm));
        dlp[c - am] = cmul(h[l], cscale(make_double2(pv * cm, pv * sm), sgn));
      }
    }
  }
}

// Per-workgroup LDS set-up shared by the assembly kernels: Gaunt and i^(l-l')
// tables, Pbar recursion coefficients and l(k).
__device__ __forceinline__ void load_tables(const double *__restrict__ cgnt_g,
                                            const double2 *__restrict__ illp_g, int lmax_tab,
                                            double *cg, double2 *il, double (*pa)[kLmaxD + 1],
                                            double (*pb)[kLmaxD + 1], double *pmmf, double *pm1f,
                                            int *lofk) {
  const int tid = threadIdx.x;
  const int kkr_tab = (lmax_tab + 1) * (lmax_tab + 1);
  for (int t = tid; t < (lmax_tab + 1) * kkr_tab * kkr_tab; t += 256) cg[t] = cgnt_g[t];
  for (int t = tid; t < kkr_tab * kkr_tab; t += 256) il[t] = illp_g[t];
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
}

// t of site i packed as the F77 route does (calculateTauMatrix.cpp:155-183:
// the leading kkri block of each spin block), each spin block placed at a
// 16-aligned offset and zero-padded for the MFMA products.
__device__ __forceinline__ void load_tpack(const double2 *__restrict__ T, int nsc, int kkrsz,
                                           int ldt, int kkri, double2 (*tn)[2 * kMaxKkr + 1]) {
  for (int t = threadIdx.x; t < 4 * kMaxKkr * kMaxKkr; t += 256) {
    int r = t % (2 * kMaxKkr), c = t / (2 * kMaxKkr);
    int rs = r / kMaxKkr, ri = r % kMaxKkr, cs = c / kMaxKkr, ci = c % kMaxKkr;
    double2 v = make_double2(0.0, 0.0);
    if (rs < nsc && cs < nsc && ri < kkri && ci < kkri)
      v = T[(rs * kkrsz + ri) + (long)ldt * (cs * kkrsz + ci)];
    tn[r][c] = v;
  }
}

// G(L2, L1) = 4pi i^(l2-l1) sum_l3 cgnt(l3/2, L1, L2) D_{l3, m2-m1}
// (makegij_c.f:240-263) for the wave's current pair, 16 x 16 zero-padded,
// G(lm2, lm1) at g[lm2 + 16 lm1].
__device__ __forceinline__ void gaunt_sum(const double *cg, const double2 *il, const int *lofk,
                                          int lmax_tab, int kkri, int kkrj, bool prel0,
                                          const double2 *dlw, double2 *g) {
  const int lane = threadIdx.x & 63;
  const int kkr_tab = (lmax_tab + 1) * (lmax_tab + 1);
  const double pi4 = 4.0 * 3.141592653589793238462643383279502884;
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
        const double2 dv = dlw[l3 * (l3 + 1) + m3];
        sacc.x += c * dv.x;
        sacc.y += c * dv.y;
      }
      gv = cscale(cmul(il[lm2 + kkr_tab * lm1], sacc), pi4);
    }
    g[e] = gv;
  }
}

// The four entries G(lm2 = 4u + kq, lm1 = li), u = 0..3, of the wave's
// current pair that lane (li, kq) feeds to the t G product on the MFMA
// (a-operand layout), by the Gaunt sum of gaunt_sum, in registers, over the
// whole 16 x 16 range (the entries outside a block's sites meet zero t
// entries or are not stored).  dlw holds all 49 D_{l,m}.
__device__ __forceinline__ void gaunt_lane(const double *cg, const double2 *il, const int *lofk,
                                           int lmax_tab, bool prel0, const double2 *dlw,
                                           double2 (&g)[4]) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  const int kkr_tab = (lmax_tab + 1) * (lmax_tab + 1);
  const double pi4 = 4.0 * 3.141592653589793238462643383279502884;
  const int lm1 = li;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int lm2 = 4 * u + kq;
    double2 gv = make_double2(0.0, 0.0);
    if (lm2 < kkr_tab && lm1 < kkr_tab) {
      const int l1 = lofk[lm1], l2 = lofk[lm2];
      const int m1 = lm1 - l1 * (l1 + 1), m2 = lm2 - l2 * (l2 + 1), m3 = m2 - m1;
      const int am3 = m3 < 0 ? -m3 : m3, dl12 = l1 - l2 < 0 ? l2 - l1 : l1 - l2;
      int llow = am3 > dl12 ? am3 : dl12;
      if (prel0) llow = l1 + l2;
      double2 sacc = make_double2(0.0, 0.0);
      for (int l3 = l1 + l2; l3 >= llow; l3 -= 2) {
        const double c = cg[l3 / 2 + (lmax_tab + 1) * (lm1 + kkr_tab * lm2)];
        const double2 dv = dlw[l3 * (l3 + 1) + m3];
        sacc.x += c * dv.x;
        sacc.y += c * dv.y;
      }
      gv = cscale(cmul(il[lm2 + kkr_tab * lm1], sacc), pi4);
    }
    g[u] = gv;
  }
}

// t of site s packed as the F77 route does (calculateTauMatrix.cpp:155-183:
// the leading kkr block of each spin block) in the b-operand layout:
// tv[bi][bj][u] = t(bi kkrsz + li, bj kkrsz + 4u + kq), zero outside kkr;
// T is the site's t block (store column and energy slot), ldt its leading
// dimension.
template <int NSC>
__device__ __forceinline__ void load_t_operand(const double2 *__restrict__ T, int kkrsz, int ldt,
                                               int kkr, double2 (&tv)[NSC][NSC][4]) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int bi = 0; bi < NSC; bi++)
#pragma unroll
    for (int bj = 0; bj < NSC; bj++)
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int ci = 4 * u + kq;
        tv[bi][bj][u] = (li < kkr && ci < kkr)
                            ? T[(bi * kkrsz + li) + (long)ldt * (bj * kkrsz + ci)]
                            : make_double2(0.0, 0.0);
      }
}

// Where the assembly writes, read from the descriptor once per workgroup
// (a descriptor field read after a store is read again, and its wait would
// take in every store in flight).
struct AsmDst {
  double2 *M;    // FP64 M (or null)
  float2 *W;     // tile-major FP32 copy (or null)
  double2 *gh;   // structure constants once per site pair (or null)
  long ld;       // leading dimension of M
  int nt;        // tiles per tile column of W
};

// Block (r, c) = -tpack_r diag_spin(G) (calculateTauMatrix.cpp:240-242) on
// FP64 MFMA, 3M form, one 16 x 16 x 16 product per spin block pair (bi, bj),
// stored to M (and the tile-major FP32 copy W).  Transposed product as in
// the LU update, so register q of lane (li, kq) is element (li, kq + 4q) of
// the block and each store is a run of 16 consecutive rows.
template <int NSC, bool WRITE_W>
__device__ __forceinline__ void store_block(const AsmDst &o, const double2 (&g)[4],
                                            const double2 (&tv)[NSC][NSC][4], int offr, int kkrr,
                                            int offc, int kkrc) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  double2 *__restrict__ M = o.M;
  float2 *__restrict__ W = o.W;
  const long ld = o.ld;
#pragma unroll
  for (int bi = 0; bi < NSC; bi++)
#pragma unroll
    for (int bj = 0; bj < NSC; bj++) {
      d4 q1 = {0.0, 0.0, 0.0, 0.0}, q2 = q1, q3 = q1;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const double2 gg = g[u], t = tv[bi][bj][u];
        q1 = __builtin_amdgcn_mfma_f64_16x16x4f64(gg.x, t.x, q1, 0, 0, 0);
        q2 = __builtin_amdgcn_mfma_f64_16x16x4f64(gg.y, t.y, q2, 0, 0, 0);
        q3 = __builtin_amdgcn_mfma_f64_16x16x4f64(gg.x + gg.y, t.x + t.y, q3, 0, 0, 0);
      }
      if (li < kkrr) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int jc = kq + 4 * q;
          if (jc < kkrc) {
            const double re = q1[q] - q2[q], im = q3[q] - q1[q] - q2[q];
            const int row = offr + bi * kkrr + li, col = offc + bj * kkrc + jc;
            // non-temporal: M (and W) are written once and read once by the
            // factorization, far beyond what the caches hold
            typedef double mv2 __attribute__((ext_vector_type(2)));
            typedef float wv2f __attribute__((ext_vector_type(2)));
            __builtin_nontemporal_store(mv2{-re, -im}, (mv2 *)&M[row + ld * col]);
            if (WRITE_W)
              __builtin_nontemporal_store(wv2f{(float)-re, (float)-im},
                                          (wv2f *)&W[tiled_index(row, col, o.nt)]);
          }
        }
      }
    }
}

typedef float f4 __attribute__((ext_vector_type(4)));

// Collinear block (r, c) = -t_r G of the FP32 copy W alone (the mixed path
// without an FP64 M): the FP32 matrix is the factorization's input, and its
// entries carry FP32 rounding either way, so the product runs on the FP32
// matrix cores, complex 4M (two accumulators, no cancellation between
// product sums).  g and tv as in store_block, rounded to FP32.  Register r
// of lane (li, kq) is element (li, 4 kq + r) of the block (the FP32 MFMA
// accumulator layout), stored as a run of 16 rows per column.
__device__ __forceinline__ void store_block_w32(const AsmDst &o, const float2 (&g)[4],
                                                const float2 (&tv)[4], int offr, int kkrr,
                                                int offc, int kkrc) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  f4 re = {0.f, 0.f, 0.f, 0.f}, im = re;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const float2 gg = g[u], t = tv[u];
    re = __builtin_amdgcn_mfma_f32_16x16x4f32(gg.x, t.x, re, 0, 0, 0);
    re = __builtin_amdgcn_mfma_f32_16x16x4f32(-gg.y, t.y, re, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f32_16x16x4f32(gg.x, t.y, im, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f32_16x16x4f32(gg.y, t.x, im, 0, 0, 0);
  }
  // non-temporal stores: W is written once here and read once by the
  // factorization, far beyond what the caches hold
  typedef float wv2 __attribute__((ext_vector_type(2)));
  if (li < kkrr) {
    float2 *__restrict__ W = o.W;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int jc = 4 * kq + r;
      if (jc < kkrc)
        __builtin_nontemporal_store(wv2{-re[r], -im[r]},
                                    (wv2 *)&W[tiled_index(offr + li, offc + jc, o.nt)]);
    }
  }
}

// t of site s in the b-operand layout of store_block_w32 (collinear), FP32;
// T and ldt as for load_t_operand.
__device__ __forceinline__ void load_t_operand_w32(const double2 *__restrict__ T, int ldt, int kkr,
                                                   float2 (&tv)[4]) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int ci = 4 * u + kq;
    const double2 v = (li < kkr && ci < kkr) ? T[li + (long)ldt * ci] : make_double2(0.0, 0.0);
    tv[u] = make_float2((float)v.x, (float)v.y);
  }
}

// The pair's structure constants once (mixed path without an FP64 M), in the
// orientation of the smaller solver position: block gh_index(p, q), p < q,
// holds G_pq (rows: the site at p), column-major.  Lane (li, kq) holds
// g[u] = G_ij(4u + kq, li) (rows: site i); G_ji = (-1)^(l1+l2) G_ij entry by
// entry (bit u of sgm set: -1), so both orientations come from the same
// entries.  Collinear,
// every site with the central site's lmax = 3 (16 x 16 blocks).
__device__ __forceinline__ void store_gh(const AsmDst &o, const double2 (&g)[4],
                                         int sgm, int i, int j, int n) {
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  const int pi = solver_pos(i, n), pj = solver_pos(j, n);
  const bool fwd = pi < pj;
  double2 *__restrict__ blk = o.gh + gh_index(fwd ? pi : pj, fwd ? pj : pi, n) * 256;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const double f = (!fwd && ((sgm >> u) & 1)) ? -1.0 : 1.0;
    blk[(4 * u + kq) + 16 * li] = make_double2(f * g[u].x, f * g[u].y);
  }
}

// Assemble for row site i = blockIdx.x of solve blockIdx.y the identity
// block (i, i) and, for the pairs j = i + k (mod n), k = 1 .. n/2 (the
// middle pair of an even n once), both blocks
//   M(i,j) = -tpack_i diag_spin(G_ij),  M(j,i) = -tpack_j diag_spin(G_ji),
// from one evaluation of D and G: D_{l,m}(-R) = (-1)^l D_{l,m}(R) and the
// Gaunt coefficients vanish unless l1 + l2 + l3 is even, so
// G_ji(L2, L1) = (-1)^(l1+l2) G_ij(L2, L1) entry by entry (exact sign flips).
// The workgroup stages the angular and Gaunt tables in LDS (16 KiB) and takes
// its pairs in chunks of at most kAsmPairs (31 KiB of D values):
//   1. D of every pair of the chunk, one lane per pair and one wave per
//      quarter of the m range, so the Hankel and Legendre recursions run once
//      per pair and |m| rather than once per (l, m) lane of a wave;
//   2. the pairs dealt round-robin to the waves: the Gaunt sums of G from
//      the chunk's D, then both blocks on the MFMA (each wave keeps the MFMA
//      operands of t_i in registers).
// NSC = n_spin_cant; MODE = kAsmM, kAsmMW or kAsmGW (the outputs written).
template <int NSC, int MODE>
__global__ __launch_bounds__(256, (MODE == kAsmGW || MODE == kAsmG) ? 3 : 1) void k_assemble(const SolveDesc *__restrict__ descs,
                                                  const double2 *__restrict__ tmat, long ldT,
                                                  const double *__restrict__ cgnt_g,
                                                  const double2 *__restrict__ illp_g,
                                                  int lmax_tab) {
  const SolveDesc &d = descs[blockIdx.y];
  const int i = blockIdx.x;
  const int n = d.numLIZ;
  if (i >= n) return;

  __shared__ double cg[(kMaxLmax + 1) * kMaxKkr * kMaxKkr];
  __shared__ double2 il[kMaxKkr * kMaxKkr];
  __shared__ double pa[kLmaxD + 1][kLmaxD + 1], pb[kLmaxD + 1][kLmaxD + 1];
  __shared__ double pmmf[kLmaxD + 1], pm1f[kLmaxD + 1];
  __shared__ int lofk[kNd];
  __shared__ double2 dl[kAsmPairs][kNd];
  __shared__ int pmeta[kAsmPairs][3];  // per pair of the chunk: kkr, block offset, store column

  const AsmDst o{d.M, d.w32, d.gh, d.ld, d.nt};
  const double2 *__restrict__ tbase = tmat + d.t_off;  // energy slot (and spin block) of the store
  const int kkrsz = d.kkrsz, ldt = d.ldt;
  const int lmi = d.lmax[i], kkri = (lmi + 1) * (lmi + 1), kkri_ns = kkri * NSC;
  const double2 *__restrict__ Ti = tbase + ldT * d.store[i];
  load_tables(cgnt_g, illp_g, lmax_tab, cg, il, pa, pb, pmmf, pm1f, lofk);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, kq = lane >> 4;
  static_assert((MODE != kAsmGW && MODE != kAsmG) || NSC == 1,
                "structure constants once per pair: collinear");
  constexpr bool gw = MODE == kAsmGW;
  constexpr bool g_only = MODE == kAsmG;
  // t_i stays in registers when collinear; the spin-canted operands (64
  // registers) are read again from L2 for each pair instead
  double2 tvi[NSC][NSC][4];
  float2 tvi32[4];
  if (gw) load_t_operand_w32(Ti, ldt, kkri, tvi32);
  else if (NSC == 1 && !g_only) load_t_operand<NSC>(Ti, kkrsz, ldt, kkri, tvi);
  __syncthreads();

  const AngTab tb{pa, pb, pmmf, pm1f};
  const double xi = d.pos[3 * i], yi = d.pos[3 * i + 1], zi = d.pos[3 * i + 2];
  const double2 prel = make_double2(d.prel_re, d.prel_im);
  const bool prel0 = (prel.x == 0.0 && prel.y == 0.0);
  const int offi = d.site_off[i];
  const int *__restrict__ lmaxp = d.lmax;
  const int *__restrict__ offp = d.site_off;
  const int *__restrict__ storep = d.store;
  const double *__restrict__ posp = d.pos;
  if (w == 0 && !g_only) {
    for (int e = lane; e < kkri_ns * kkri_ns; e += 64) {
      int rr = e % kkri_ns, cc = e / kkri_ns;
      if (MODE != kAsmGW)
        o.M[(offi + rr) + o.ld * (offi + cc)] = make_double2(rr == cc ? 1.0 : 0.0, 0.0);
      if (MODE != kAsmM)
        o.W[tiled_index(offi + rr, offi + cc, o.nt)] = make_float2(rr == cc ? 1.f : 0.f, 0.f);
    }
  }
  // (-1)^(l1+l2) of the lane's four G entries
  int sgm = 0;  // bit u: (-1)^(l1+l2) = -1 for entry u
#pragma unroll
  for (int u = 0; u < 4; u++) sgm |= ((lofk[li] + lofk[4 * u + kq]) & 1) << u;

  // the pairs j = i + k (mod n), k = 1 .. n/2, the middle pair of an even n
  // (k = n/2, the last) once; in chunks of at most kAsmPairs
  const int half = n / 2;
  const int npairs = half - ((n % 2 == 0 && (i + half) % n < i) ? 1 : 0);
  const int nchunk = (npairs + kAsmPairs - 1) / kAsmPairs;
  const int per = nchunk ? (npairs + nchunk - 1) / nchunk : 0;
  for (int q0 = 0; q0 < npairs; q0 += per) {
    const int np = min(per, npairs - q0);
    // D of the chunk's pairs: lane = pair, wave w = quarter w of the m range;
    // wave 0 also leaves each pair's site data in LDS, so that no global read
    // (whose wait would take in the stores in flight) sits between the pairs
    if (lane < np) {
      const int j = (i + q0 + lane + 1) % n;
      const double rx = xi - posp[3 * j], ry = yi - posp[3 * j + 1], rz = zi - posp[3 * j + 2];
      const double r = sqrt(rx * rx + ry * ry + rz * rz);
      if (w == 0) {
        const int lmj = lmaxp[j];
        pmeta[lane][0] = (lmj + 1) * (lmj + 1);
        pmeta[lane][1] = offp[j];
        pmeta[lane][2] = storep[j];
      }
      switch (w) {
        case 0: d_pair_part<0>(tb, prel, prel0, rx, ry, rz, r, dl[lane]); break;
        case 1: d_pair_part<1>(tb, prel, prel0, rx, ry, rz, r, dl[lane]); break;
        case 2: d_pair_part<2>(tb, prel, prel0, rx, ry, rz, r, dl[lane]); break;
        default: d_pair_part<3>(tb, prel, prel0, rx, ry, rz, r, dl[lane]); break;
      }
    }
    __syncthreads();
    // G and both blocks of each pair, pairs dealt round-robin to the waves
    if (g_only) {
      for (int pq = w; pq < np; pq += 4) {
        double2 g[4];
        gaunt_lane(cg, il, lofk, lmax_tab, prel0, dl[pq], g);
        store_gh(o, g, sgm, i, (i + q0 + pq + 1) % n, n);
      }
    } else if (gw) {
      for (int pq = w; pq < np; pq += 4) {
        const int j = (i + q0 + pq + 1) % n;
        const int kkrj = pmeta[pq][0], offj = pmeta[pq][1];
        float2 tvj32[4];
        load_t_operand_w32(tbase + ldT * pmeta[pq][2], ldt, kkrj, tvj32);
        double2 g[4];
        gaunt_lane(cg, il, lofk, lmax_tab, prel0, dl[pq], g);
        store_gh(o, g, sgm, i, j, n);
        float2 gf[4];
#pragma unroll
        for (int u = 0; u < 4; u++) gf[u] = make_float2((float)g[u].x, (float)g[u].y);
        store_block_w32(o, gf, tvi32, offi, kkri, offj, kkrj);
#pragma unroll
        for (int u = 0; u < 4; u++)
          if ((sgm >> u) & 1) gf[u] = make_float2(-gf[u].x, -gf[u].y);
        store_block_w32(o, gf, tvj32, offj, kkrj, offi, kkri);
      }
    } else {
      for (int pq = w; pq < np; pq += 4) {
        const int j = (i + q0 + pq + 1) % n;
        const int kkrj = pmeta[pq][0], offj = pmeta[pq][1];
        const double2 *__restrict__ Tj = tbase + ldT * pmeta[pq][2];
        double2 g[4];
        double2 tvj[NSC][NSC][4];
        if (NSC == 1) load_t_operand<NSC>(Tj, kkrsz, ldt, kkrj, tvj);
        else load_t_operand<NSC>(Ti, kkrsz, ldt, kkri, tvi);
        gaunt_lane(cg, il, lofk, lmax_tab, prel0, dl[pq], g);
        store_block<NSC, MODE == kAsmMW>(o, g, tvi, offi, kkri, offj, kkrj);
        if (NSC != 1) {  // spin-canted: t_j only once the block of t_i is done
          __builtin_amdgcn_sched_barrier(0);
          load_t_operand<NSC>(Tj, kkrsz, ldt, kkrj, tvj);
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
          if ((sgm >> u) & 1) g[u] = make_double2(-g[u].x, -g[u].y);
        store_block<NSC, MODE == kAsmMW>(o, g, tvj, offj, kkrj, offi, kkri);
      }
    }
    __syncthreads();  // dl and pmeta are rewritten by the next chunk
  }
}

#define LSMSX_ASM_INST(NSC, MODE)                                                        \
  template __global__ void k_assemble<NSC, MODE>(const SolveDesc *, const double2 *, long, \
                                                 const double *, const double2 *, int);
LSMSX_ASM_INST(1, kAsmM)
LSMSX_ASM_INST(1, kAsmMW)
LSMSX_ASM_INST(1, kAsmGW)
LSMSX_ASM_INST(1, kAsmG)
LSMSX_ASM_INST(2, kAsmM)
LSMSX_ASM_INST(2, kAsmMW)
#undef LSMSX_ASM_INST

// The blocks of one solve from structure constants already stored once per
// site pair (d.gh, written by k_assemble<1, kAsmG> for the first solve of a
// group that shares the LIZ geometry and the energy): for row site i =
// blockIdx.x of solve blockIdx.y, the identity block (i, i) and, for the
// pairs j = i + k (mod n) of k_assemble, both blocks
//   M(i,j) = -t_i G_ij,  M(j,i) = -t_j G_ji,  G_ji = (-1)^(l1+l2) G_ij,
// with G_ij read back in the orientation store_gh wrote it (the same FP64
// values the solve would compute, so the blocks are bitwise those of
// k_assemble).  Collinear; MODE = kAsmM, kAsmMW or kAsmGW (W alone).
template <int MODE>
__global__ __launch_bounds__(256) void k_assemble_from_g(const SolveDesc *__restrict__ descs,
                                                         const double2 *__restrict__ tmat,
                                                         long ldT) {
  const SolveDesc &d = descs[blockIdx.y];
  const int i = blockIdx.x;
  const int n = d.numLIZ;
  if (i >= n) return;
  const AsmDst o{d.M, d.w32, d.gh, d.ld, d.nt};
  const double2 *__restrict__ tbase = tmat + d.t_off;
  const int kkrsz = d.kkrsz, ldt = d.ldt;
  const int *__restrict__ lmaxp = d.lmax;
  const int *__restrict__ offp = d.site_off;
  const int *__restrict__ storep = d.store;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, kq = lane >> 4;
  const int lmi = lmaxp[i], kkri = (lmi + 1) * (lmi + 1);
  const int offi = offp[i];
  const double2 *__restrict__ Ti = tbase + ldT * storep[i];
  constexpr bool w_only = MODE == kAsmGW;
  double2 tvi[1][1][4];
  float2 tvi32[4];
  if (w_only) load_t_operand_w32(Ti, ldt, kkri, tvi32);
  else load_t_operand<1>(Ti, kkrsz, ldt, kkri, tvi);
  if (w == 0) {
    for (int e = lane; e < kkri * kkri; e += 64) {
      const int rr = e % kkri, cc = e / kkri;
      if (MODE != kAsmGW) o.M[(offi + rr) + o.ld * (offi + cc)] = make_double2(rr == cc ? 1.0 : 0.0, 0.0);
      if (MODE != kAsmM) o.W[tiled_index(offi + rr, offi + cc, o.nt)] = make_float2(rr == cc ? 1.f : 0.f, 0.f);
    }
  }
  // (-1)^(l1+l2) of the lane's four G entries, l of an lm index = floor(sqrt(lm))
  auto lof = [](int lm) { return lm < 1 ? 0 : lm < 4 ? 1 : lm < 9 ? 2 : 3; };
  int sgm = 0;
#pragma unroll
  for (int u = 0; u < 4; u++) sgm |= ((lof(li) + lof(4 * u + kq)) & 1) << u;
  const int pi = solver_pos(i, n);
  const int half = n / 2;
  const int npairs = half - ((n % 2 == 0 && (i + half) % n < i) ? 1 : 0);
  for (int k = 1 + w; k <= npairs; k += 4) {
    const int j = (i + k) % n;
    const int lmj = lmaxp[j], kkrj = (lmj + 1) * (lmj + 1), offj = offp[j];
    const double2 *__restrict__ Tj = tbase + ldT * storep[j];
    const int pj = solver_pos(j, n);
    const bool fwd = pi < pj;
    const double2 *__restrict__ blk = o.gh + gh_index(fwd ? pi : pj, fwd ? pj : pi, n) * 256;
    double2 g[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const double2 v = blk[(4 * u + kq) + 16 * li];
      g[u] = (!fwd && ((sgm >> u) & 1)) ? make_double2(-v.x, -v.y) : v;
    }
    if (w_only) {
      float2 tvj32[4], gf[4];
      load_t_operand_w32(Tj, ldt, kkrj, tvj32);
#pragma unroll
      for (int u = 0; u < 4; u++) gf[u] = make_float2((float)g[u].x, (float)g[u].y);
      store_block_w32(o, gf, tvi32, offi, kkri, offj, kkrj);
#pragma unroll
      for (int u = 0; u < 4; u++)
        if ((sgm >> u) & 1) gf[u] = make_float2(-gf[u].x, -gf[u].y);
      store_block_w32(o, gf, tvj32, offj, kkrj, offi, kkri);
    } else {
      double2 tvj[1][1][4];
      load_t_operand<1>(Tj, kkrsz, ldt, kkrj, tvj);
      store_block<1, MODE == kAsmMW>(o, g, tvi, offi, kkri, offj, kkrj);
#pragma unroll
      for (int u = 0; u < 4; u++)
        if ((sgm >> u) & 1) g[u] = make_double2(-g[u].x, -g[u].y);
      store_block<1, MODE == kAsmMW>(o, g, tvj, offj, kkrj, offi, kkri);
    }
  }
}

template __global__ void k_assemble_from_g<kAsmM>(const SolveDesc *, const double2 *, long);
template __global__ void k_assemble_from_g<kAsmMW>(const SolveDesc *, const double2 *, long);
template __global__ void k_assemble_from_g<kAsmGW>(const SolveDesc *, const double2 *, long);

// Energy derivative of the assembly contracted with the explicit inverse, one
// workgroup per row site i = blockIdx.x of solve blockIdx.y:
//   lpart[i] = sum_j Tr(Minv_ji dM_ij/dE),
//     dM_ij/dE = -(tdot_i diag_spin(G_ij) + t_i diag_spin(dG_ij/dprel) dprel/dE)
//   gx(site i rows) = sum_j diag_spin(G_ij) X_j,   X = M^-1 [0; t0]  (d.X)
// tdot = dt/dE from the derivative store dtmat (same layout as tmat); with
// dtmat == nullptr only gx is formed.  The (i, i) blocks are constant (M_ii =
// 1, G_ii = 0) and contribute nothing.
__global__ __launch_bounds__(256) void k_lloyd(const SolveDesc *__restrict__ descs,
                                               const double2 *__restrict__ tmat,
                                               const double2 *__restrict__ dtmat, long ldT,
                                               const double *__restrict__ cgnt_g,
                                               const double2 *__restrict__ illp_g,
                                               int lmax_tab) {
  const SolveDesc &d = descs[blockIdx.y];
  const int i = blockIdx.x;
  if (i >= d.numLIZ) return;

  __shared__ double cg[(kMaxLmax + 1) * kMaxKkr * kMaxKkr];
  __shared__ double2 il[kMaxKkr * kMaxKkr];
  __shared__ double2 tn[2 * kMaxKkr][2 * kMaxKkr + 1];
  __shared__ double2 tdn[2 * kMaxKkr][2 * kMaxKkr + 1];
  __shared__ double pa[kLmaxD + 1][kLmaxD + 1], pb[kLmaxD + 1][kLmaxD + 1];
  __shared__ double pmmf[kLmaxD + 1], pm1f[kLmaxD + 1];
  __shared__ int lofk[kNd];
  __shared__ double2 dl[4][kNd], ddl[4][kNd];
  __shared__ double2 gs[4][kMaxKkr * kMaxKkr], gds[4][kMaxKkr * kMaxKkr];
  __shared__ double2 red[2 * kMaxKkr][2 * kMaxKkr + 1];
  __shared__ double2 trs[4];

  const int nsc = d.n_spin_cant, kc = d.kc, Ns = d.Ns;
  const int lmi = d.lmax[i], kkri = (lmi + 1) * (lmi + 1), kkri_ns = kkri * nsc;
  const bool want_tr = dtmat != nullptr;
  load_tables(cgnt_g, illp_g, lmax_tab, cg, il, pa, pb, pmmf, pm1f, lofk);
  load_tpack(tmat + ldT * d.store[i] + d.t_off, nsc, d.kkrsz, d.ldt, kkri, tn);
  if (want_tr) load_tpack(dtmat + ldT * d.store[i] + d.t_off, nsc, d.kkrsz, d.ldt, kkri, tdn);
  __syncthreads();

  const AngTab tb{pa, pb, pmmf, pm1f};
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  const double xi = d.pos[3 * i], yi = d.pos[3 * i + 1], zi = d.pos[3 * i + 2];
  const double2 prel = make_double2(d.prel_re, d.prel_im);
  const double2 dprel = make_double2(d.dprel_re, d.dprel_im);
  const long ld = d.ld;
  const int offi = d.site_off[i];
  const double2 *__restrict__ Minv = d.Minv;
  const double2 *__restrict__ X = d.X;
  const int ncb = kc / 16;
  double2 tr = make_double2(0.0, 0.0);
  d4 gq[2][2][3];
#pragma unroll
  for (int bs = 0; bs < 2; bs++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int m = 0; m < 3; m++) gq[bs][cb][m] = d4{0.0, 0.0, 0.0, 0.0};

  for (int j = w; j < d.numLIZ; j += 4) {
    if (j == i) continue;
    const int lmj = d.lmax[j], kkrj = (lmj + 1) * (lmj + 1);
    const int offj = d.site_off[j];
    const double rx = xi - d.pos[3 * j], ry = yi - d.pos[3 * j + 1], rz = zi - d.pos[3 * j + 2];
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    const int lend = lmi + lmj;
    if (lane < (lend + 1) * (lend + 1)) {
      const int l = lofk[lane], m = lane - l * (l + 1);
      double2 dh;
      const double2 h = hankel_lane(prel, r, l, &dh);
      const double2 ang = angular_lane(tb, l, m, rx, ry, rz, r);
      dl[w][lane] = cmul(h, ang);
      ddl[w][lane] = cmul(dprel, cmul(dh, ang));
    }
    wave_sync();
    gaunt_sum(cg, il, lofk, lmax_tab, kkri, kkrj, false, dl[w], gs[w]);
    if (want_tr) gaunt_sum(cg, il, lofk, lmax_tab, kkri, kkrj, false, ddl[w], gds[w]);
    wave_sync();
    // trace part: D = -(tdot G + t G'), element (li, kq + 4r) of spin block
    // (bi, bj), against Minv(site j row bj*kkrj + kq + 4r, site i col bi*kkri + li)
    if (want_tr) {
      for (int bi = 0; bi < nsc; bi++)
        for (int bj = 0; bj < nsc; bj++) {
          d4 q1 = {0.0, 0.0, 0.0, 0.0}, q2 = q1, q3 = q1;
#pragma unroll
          for (int k0 = 0; k0 < kMaxKkr; k0 += 4) {
            const double2 g = gs[w][(k0 + kq) + kMaxKkr * li];
            const double2 t = tdn[bi * kMaxKkr + li][bj * kMaxKkr + k0 + kq];
            q1 = __builtin_amdgcn_mfma_f64_16x16x4f64(g.x, t.x, q1, 0, 0, 0);
            q2 = __builtin_amdgcn_mfma_f64_16x16x4f64(g.y, t.y, q2, 0, 0, 0);
            q3 = __builtin_amdgcn_mfma_f64_16x16x4f64(g.x + g.y, t.x + t.y, q3, 0, 0, 0);
          }
#pragma unroll
          for (int k0 = 0; k0 < kMaxKkr; k0 += 4) {
            const double2 g = gds[w][(k0 + kq) + kMaxKkr * li];
            const double2 t = tn[bi * kMaxKkr + li][bj * kMaxKkr + k0 + kq];
            q1 = __builtin_amdgcn_mfma_f64_16x16x4f64(g.x, t.x, q1, 0, 0, 0);
            q2 = __builtin_amdgcn_mfma_f64_16x16x4f64(g.y, t.y, q2, 0, 0, 0);
            q3 = __builtin_amdgcn_mfma_f64_16x16x4f64(g.x + g.y, t.x + t.y, q3, 0, 0, 0);
          }
          if (li < kkri) {
#pragma unroll
            for (int r4 = 0; r4 < 4; r4++) {
              const int jc = kq + 4 * r4;
              if (jc < kkrj) {
                const double2 dv = make_double2(-(q1[r4] - q2[r4]), -(q3[r4] - q1[r4] - q2[r4]));
                const double2 mv = Minv[(offj + bj * kkrj + jc) + ld * (offi + bi * kkri + li)];
                tr.x += mv.x * dv.x - mv.y * dv.y;
                tr.y += mv.x * dv.y + mv.y * dv.x;
              }
            }
          }
        }
    }
    // (G X)_i part: spin block bs rows li, column 16 cb + kq + 4r
#pragma unroll
    for (int bs = 0; bs < 2; bs++)
#pragma unroll
      for (int cb = 0; cb < 2; cb++) {
        if (bs >= nsc || cb >= ncb) continue;
#pragma unroll
        for (int k0 = 0; k0 < kMaxKkr; k0 += 4) {
          const int kk = k0 + kq;
          double2 x = make_double2(0.0, 0.0);
          if (kk < kkrj) x = X[(offj + bs * kkrj + kk) + (long)Ns * (16 * cb + li)];
          const double2 g = gs[w][li + kMaxKkr * kk];
          gq[bs][cb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, g.x, gq[bs][cb][0], 0, 0, 0);
          gq[bs][cb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, g.y, gq[bs][cb][1], 0, 0, 0);
          gq[bs][cb][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x + x.y, g.x + g.y,
                                                               gq[bs][cb][2], 0, 0, 0);
        }
      }
    wave_sync();
  }
  // reductions: trace over lanes and waves, (G X)_i over waves in wave order
  for (int off = 32; off >= 1; off >>= 1) {
    tr.x += __shfl_xor(tr.x, off, 64);
    tr.y += __shfl_xor(tr.y, off, 64);
  }
  if (lane == 0) trs[w] = tr;
  for (int ww = 0; ww < 4; ww++) {
    if (w == ww) {
#pragma unroll
      for (int bs = 0; bs < 2; bs++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
          for (int r4 = 0; r4 < 4; r4++) {
            if (bs >= nsc || cb >= ncb) continue;
            const d4 *q = gq[bs][cb];
            const double2 v = make_double2(q[0][r4] - q[1][r4], q[2][r4] - q[0][r4] - q[1][r4]);
            double2 &o = red[bs * kMaxKkr + li][16 * cb + kq + 4 * r4];
            if (ww == 0) o = v;
            else o = make_double2(o.x + v.x, o.y + v.y);
          }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double2 s = make_double2(0.0, 0.0);
    for (int ww = 0; ww < 4; ww++) {
      s.x += trs[ww].x;
      s.y += trs[ww].y;
    }
    d.lpart[i] = s;
  }
  for (int e = threadIdx.x; e < kkri_ns * kc; e += 256) {
    const int p = e % kkri_ns, col = e / kkri_ns;
    const int bs = p / kkri, a = p % kkri;
    d.gx[(offi + p) + (long)Ns * col] = red[bs * kMaxKkr + a][col];
  }
}

// Identity in the leading pad rows/columns of the solver layout.
__global__ __launch_bounds__(256) void k_pad_identity(const SolveDesc *__restrict__ descs) {
  const SolveDesc &d = descs[blockIdx.x];
  const int pad = d.pad;
  if (pad == 0) return;
  double2 *M = d.M;
  float2 *W = d.w32;
  const long ld = d.ld;
  // columns 0..pad-1: full columns; columns pad..Ns-1: rows 0..pad-1
  for (long e = threadIdx.x; e < (long)pad * d.Ns; e += blockDim.x) {
    int rr = (int)(e % d.Ns), cc = (int)(e / d.Ns);
    if (M) M[rr + ld * cc] = make_double2(rr == cc ? 1.0 : 0.0, 0.0);
    if (W) W[tiled_index(rr, cc, d.nt)] = make_float2(rr == cc ? 1.f : 0.f, 0.f);
  }
  for (long e = threadIdx.x; e < (long)pad * (d.Ns - pad); e += blockDim.x) {
    int rr = (int)(e % pad), cc = pad + (int)(e / pad);
    if (M) M[rr + ld * cc] = make_double2(0.0, 0.0);
    if (W) W[tiled_index(rr, cc, d.nt)] = make_float2(0.f, 0.f);
  }
}

}  // namespace lsmsx
