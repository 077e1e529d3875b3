/*
 * lsms_oracle.c -- CPU restatement of the LSMS tau00 path (KKR assembly plus
 * central-block inversion), written from the reference's algorithm.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product library never links or calls it.
 *
 * Every function names the reference file:line it restates.  How each one is
 * pinned (golden values, the compiled reference routines in oracle/_ref,
 * closed forms, LAPACK) is stated in the tests that exercise it.
 */
#include "lsms_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* include/PhysicalConstants.hpp:5-8 */
double or_cphot(void) { return 2.0 * 137.035999206; }

/* calculateTauMatrix.cpp:846: prel = sqrt(E (1 + E c^-2)) */
zc or_prel(zc e) {
  double c2inv = 1.0 / (or_cphot() * or_cphot());
  return csqrt(e * (1.0 + e * c2inv));
}

/* Misc/Indices.hpp:27-61: k = l(l+1)+m, l = 0..lmax, m = -l..l */
void or_lofk_mofk(int lmax, int *lofk, int *mofk) {
  int k = 0;
  for (int l = 0; l <= lmax; l++)
    for (int m = -l; m <= l; m++) {
      lofk[k] = l;
      mofk[k] = m;
      k++;
    }
}

/*
 * Gauss-Legendre nodes and weights on [x1,x2] by Newton iteration on P_n,
 * nodes ascending (Misc/calculateGaussLegendrePoints.hpp:13-76).
 */
void or_gauss_legendre(int n, double x1, double x2, double *x, double *w) {
  const double pi = acos(-1.0);
  double half_w = 0.5 * (x2 - x1), mid = 0.5 * (x2 + x1);
  for (int i = 0; i < n; i++) {
    double xi = cos(pi * ((double)(i + 1) - 0.25) / ((double)n + 0.5));
    double dx = 1000.0;
    double pk = 0, pkm1 = 0, pkp1 = 0, ppn = 0;
    while (xi + dx / 4.0 != xi) {
      pk = 0.0;
      pkp1 = 1.0;
      for (int k = 0; k <= n; k++) {
        pkm1 = pk;
        pk = pkp1;
        pkp1 = ((double)(2 * k + 1) * xi * pk - (double)k * pkm1) / (double)(k + 1);
      }
      ppn = (double)n * (xi * pk - pkm1) / (xi * xi - 1.0);
      dx = pk / ppn;
      xi = xi - dx;
    }
    pk = 0.0;
    pkp1 = 1.0;
    for (int k = 0; k <= n; k++) {
      pkm1 = pk;
      pk = pkp1;
      pkp1 = ((double)(2 * k + 1) * xi * pk - (double)k * pkm1) / (double)(k + 1);
    }
    ppn = (double)n * (xi * pk - pkm1) / (xi * xi - 1.0);
    w[i] = half_w * (-2.0 / ((double)(n + 1) * ppn * pkp1));
    x[i] = mid - half_w * xi;
  }
}

/*
 * Normalized associated Legendre functions, Condon-Shortley phase included,
 * index l(l+1)/2+m (Misc/associatedLegendreFunction.hpp:33-69).
 */
void or_plm_normalized(int lmax, double x, double *plm) {
  const double pi = acos(-1.0);
  double y = sqrt(1.0 - x * x);
#define PIDX(l, m) ((l) * ((l) + 1) / 2 + (m))
  plm[0] = sqrt(1.0 / (4.0 * pi));
  if (lmax < 1) return;
  for (int m = 1; m <= lmax; m++) {
    plm[PIDX(m, m)] =
        -sqrt((double)(2 * m + 1) / (double)(2 * m)) * y * plm[PIDX(m - 1, m - 1)];
    plm[PIDX(m, m - 1)] = sqrt((double)(2 * m + 1)) * x * plm[PIDX(m - 1, m - 1)];
  }
  for (int m = 0; m < lmax; m++)
    for (int l = m + 2; l <= lmax; l++) {
      double a = sqrt((double)(4 * l * l - 1) / (double)(l * l - m * m));
      double b = sqrt((double)((l - 1) * (l - 1) - m * m) /
                      (double)(4 * (l - 1) * (l - 1) - 1));
      plm[PIDX(l, m)] = a * (x * plm[PIDX(l - 1, m)] - b * plm[PIDX(l - 2, m)]);
    }
#undef PIDX
}

/* Misc/gaunt.f defac: index of P_l|m| and the (-1)^m factor for m < 0 */
static void defac(int l, int m, int *ifac, int *jl) {
  if (m >= 0) {
    *ifac = 1;
    *jl = l * (l + 1) / 2 + m;
  } else {
    *ifac = ((-m) % 2) ? -1 : 1;
    *jl = l * (l + 1) / 2 - m;
  }
}

/*
 * Gaunt numbers cgnt(l3/2, j1, j2) = 4pi * sum_g w_g P1 P2 P3 over the first
 * 2lmax+1 nodes of a 2(2lmax+1)-point rule, zeroed below 1e-14
 * (Misc/cgaunt_c.f90:55-101, Misc/gaunt.f).  cgnt has dims (lmax+1,kkr,kkr).
 */
void or_cgaunt(int lmax, double *cgnt) {
  int lm2 = 2 * lmax;
  int ng_all = 2 * (2 * lmax + 1), ng = 2 * lmax + 1;
  int ndlm2 = (lm2 + 1) * (lm2 + 2) / 2;
  int kkr = (lmax + 1) * (lmax + 1);
  double *tg = malloc(sizeof(double) * ng_all), *wg = malloc(sizeof(double) * ng_all);
  double *plmg = malloc(sizeof(double) * ndlm2 * ng);
  int *lofk = malloc(sizeof(int) * kkr), *mofk = malloc(sizeof(int) * kkr);
  const double fourpi = 4.0 * 3.141592653589793;
  or_gauss_legendre(ng_all, -1.0, 1.0, tg, wg);
  for (int g = 0; g < ng; g++) or_plm_normalized(lm2, tg[g], plmg + (size_t)g * ndlm2);
  or_lofk_mofk(lmax, lofk, mofk);
  memset(cgnt, 0, sizeof(double) * (lmax + 1) * kkr * kkr);
  for (int j2 = 0; j2 < kkr; j2++) {
    int l2 = lofk[j2], m2 = mofk[j2];
    for (int j1 = 0; j1 < kkr; j1++) {
      int l1 = lofk[j1], m1 = mofk[j1];
      int m3 = m2 - m1;
      int lo = abs(m3) > abs(l1 - l2) ? abs(m3) : abs(l1 - l2);
      for (int l3 = l1 + l2; l3 >= lo; l3 -= 2) {
        double val = 0.0;
        /* gaunt(l1,m1,l2,m2,l3,m3): parity, m and triangle selection rules */
        if (((l1 + l2 + l3) % 2) == 0 && (m1 - m2 + m3) == 0 &&
            !(l1 + l2 < l3 || l2 + l3 < l1 || l3 + l1 < l2)) {
          int if1, j1l, if2, j2l, if3, j3l;
          defac(l1, m1, &if1, &j1l);
          defac(l2, -m2, &if2, &j2l);
          defac(l3, m3, &if3, &j3l);
          double s = 0.0;
          for (int g = 0; g < ng; g++)
            s += wg[g] * plmg[j1l + g * ndlm2] * plmg[j2l + g * ndlm2] *
                 plmg[j3l + g * ndlm2];
          if (fabs(s) < 1.0e-14)
            val = 0.0;
          else
            val = fourpi * s * if1 * ((abs(m2) % 2) ? -1 : 1) * if2 * if3;
        }
        if (fabs(val) < 1.0e-14) val = 0.0;
        cgnt[l3 / 2 + (lmax + 1) * (j1 + kkr * j2)] = val;
      }
    }
  }
  free(tg); free(wg); free(plmg); free(lofk); free(mofk);
}

static zc ipow(int n) { /* i^n, exact */
  switch (((n % 4) + 4) % 4) {
    case 0: return 1.0;
    case 1: return I;
    case 2: return -1.0;
    default: return -I;
  }
}

/*
 * illp(lm,lm') = i^(l-l'), ilp1(l) = i^(l+1) for l = 0..2lmax
 * (Misc/ifacts_c.f:31-58).  illp is kkr x kkr column-major.
 */
void or_ifacts(int lmax, zc *illp, zc *ilp1) {
  int kkr = (lmax + 1) * (lmax + 1);
  for (int lm = 0; lm < kkr; lm++) {
    int l = (int)floor(sqrt((double)lm) + 1e-9);
    for (int lmp = 0; lmp < kkr; lmp++) {
      int lp = (int)floor(sqrt((double)lmp) + 1e-9);
      illp[lm + kkr * lmp] = ipow(l - lp);
    }
  }
  for (int l = 0; l <= 2 * lmax; l++) ilp1[l] = ipow(l + 1);
}

/*
 * hfn_l = -i^(l+1) kappa h_l(kappa r) by upward recursion of
 * p_l = z e^{-iz} h_l(z); prel = 0 uses the Coulomb (static) limit
 * (MultipleScattering/makegij_c.f:139-163; buildKKRMatrix_CPU.cpp:23-40).
 */
void or_hankel(zc prel, double r, int lend, zc *hfn) {
  if (prel == 0.0) {
    hfn[0] = 1.0 / r;
    for (int l = 1; l <= lend; l++) hfn[l] = I * (double)(2 * l - 1) * hfn[l - 1] / r;
    return;
  }
  zc z = prel * r;
  hfn[0] = -I;
  hfn[1] = -1.0 - I / z;
  for (int l = 1; l < lend; l++) hfn[l + 1] = (double)(2 * l + 1) * hfn[l] / z - hfn[l - 1];
  zc f = cexp(I * z) / r;
  for (int l = 0; l <= lend; l++) hfn[l] = -hfn[l] * f * ipow(l + 1);
}

/*
 * Free-electron structure constant block G_ij (kkri x kkrj) for
 * rij = R_i - R_j (MultipleScattering/makegij_c.f:100-263).
 *   G(lm2,lm1) = 4pi i^(l2-l1) sum_l3 cgnt(l3/2,lm1,lm2) D(l3, m2-m1)
 * with D(l,m) = hfn_l * Ybar-type angular factor.
 * cgnt has dims (lmax_tab+1, kkr_tab, kkr_tab); illp is kkr_tab square.
 */
void or_makegij(int lmaxi, int lmaxj, int lmax_tab, zc prel, const double *rij,
                const double *cgnt, const zc *illp, const zc *ilp1, zc *gij) {
  (void)ilp1;
  int lend = lmaxi + lmaxj;
  int kkri = (lmaxi + 1) * (lmaxi + 1), kkrj = (lmaxj + 1) * (lmaxj + 1);
  int kkr_tab = (lmax_tab + 1) * (lmax_tab + 1);
  zc hfn[2 * 16 + 2];
  double plm[(2 * 16 + 1) * (2 * 16 + 2) / 2];
  double cosmp[2 * 16 + 2], sinmp[2 * 16 + 2];
  zc dlm[(2 * 16 + 1) * (2 * 16 + 1)];
  double rmag = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
  const double pi4 = 4.0 * 2.0 * asin(1.0);

  or_hankel(prel, rmag, lend > 0 ? lend : 1, hfn);
  or_plm_normalized(lend, rij[2] / rmag, plm);

  double pmag = sqrt(rij[0] * rij[0] + rij[1] * rij[1]);
  cosmp[0] = 1.0;
  sinmp[0] = 0.0;
  if (pmag >= 1.0e-6) {
    cosmp[1] = rij[0] / pmag;
    sinmp[1] = rij[1] / pmag;
  } else {
    cosmp[1] = 0.0;
    sinmp[1] = 0.0;
  }
  for (int m = 2; m <= lend; m++) {
    cosmp[m] = cosmp[m - 1] * cosmp[1] - sinmp[m - 1] * sinmp[1];
    sinmp[m] = sinmp[m - 1] * cosmp[1] + cosmp[m - 1] * sinmp[1];
  }
  for (int l = 0; l <= lend; l++) {
    int j = l * (l + 1), ll = j / 2;
    double m1m = 1.0;
    dlm[j] = hfn[l] * plm[ll];
    for (int m = 1; m <= l; m++) {
      m1m = -m1m;
      zc fac = plm[ll + m] * (cosmp[m] + I * sinmp[m]);
      dlm[j - m] = hfn[l] * m1m * fac;
      dlm[j + m] = hfn[l] * conj(fac);
    }
  }
  for (int lm1 = 0; lm1 < kkrj; lm1++) {
    int l1 = (int)floor(sqrt((double)lm1) + 1e-9), m1 = lm1 - l1 * (l1 + 1);
    for (int lm2 = 0; lm2 < kkri; lm2++) {
      int l2 = (int)floor(sqrt((double)lm2) + 1e-9), m2 = lm2 - l2 * (l2 + 1);
      int m3 = m2 - m1;
      int llow = abs(m3) > abs(l1 - l2) ? abs(m3) : abs(l1 - l2);
      if (prel == 0.0) llow = l1 + l2;
      zc s = 0.0;
      for (int l3 = l1 + l2; l3 >= llow; l3 -= 2)
        s += cgnt[l3 / 2 + (lmax_tab + 1) * (lm1 + kkr_tab * lm2)] *
             dlm[l3 * (l3 + 1) + m3];
      gij[lm2 + kkri * lm1] = pi4 * illp[lm2 + kkr_tab * lm1] * s;
    }
  }
}

/* nrmat = sum_i (LIZlmax_i+1)^2 (buildLIZandCommLists.cpp:393-399) */
int or_nrmat(int numLIZ, const int *LIZlmax) {
  int n = 0;
  for (int i = 0; i < numLIZ; i++) n += (LIZlmax[i] + 1) * (LIZlmax[i] + 1);
  return n;
}

/*
 * Pack the kkr1_ns x kkr1_ns t-matrix of a site from tmatStore
 * (calculateTauMatrix.cpp:155-183): the (is,js) spin blocks keep their
 * leading kkr1 x kkr1 part; the collinear case takes the ispin block.
 */
static void pack_tmat(int kkrsz, int kkr1, int n_spin_cant, int n_spin_pola,
                      int ispin, const zc *tcol, long blkoff, zc *tmat_n) {
  int kkrsz_ns = kkrsz * n_spin_cant, kkr1_ns = kkr1 * n_spin_cant;
  if (n_spin_pola == n_spin_cant) {
    for (int js = 0; js < n_spin_cant; js++)
      for (int j = 0; j < kkr1; j++)
        for (int is = 0; is < n_spin_cant; is++)
          for (int i = 0; i < kkr1; i++)
            tmat_n[(is * kkr1 + i) + kkr1_ns * (js * kkr1 + j)] =
                tcol[blkoff + (kkrsz * is + i) + (long)kkrsz_ns * (kkrsz * js + j)];
  } else {
    long jsm = (long)kkrsz * kkrsz * ispin;
    for (int j = 0; j < kkr1; j++)
      for (int i = 0; i < kkr1; i++)
        tmat_n[i + kkr1 * j] = tcol[blkoff + jsm + i + (long)kkrsz_ns * j];
  }
}

/*
 * M = 1 - t_i G_ij over the LIZ, F77 route
 * (calculateTauMatrix.cpp:57-315; makegij_c.f; setgij.f:29-38).
 * LIZPos is 3 x numLIZ column-major, relative to the central site (site 0).
 * tmatStore column s, energy slot iie: tmatStore[iie*blk + r + s*ldTmat].
 * Returns nrmat_ns.
 */
int or_build_kkr(int maxlmax, int n_spin_cant, int n_spin_pola, int ispin,
                 int numLIZ, const double *LIZPos, const int *LIZlmax,
                 const int *LIZStoreIdx, const zc *tmatStore, long ldTmatStore,
                 long blkSizeTmatStore, int iie, zc prel, zc *m, int ldm) {
  int kkrsz = (maxlmax + 1) * (maxlmax + 1);
  int nrmat_ns = n_spin_cant * or_nrmat(numLIZ, LIZlmax);
  int kkr_tab = kkrsz;
  double *cgnt = malloc(sizeof(double) * (maxlmax + 1) * kkr_tab * kkr_tab);
  zc *illp = malloc(sizeof(zc) * kkr_tab * kkr_tab);
  zc ilp1[64];
  zc *gij = malloc(sizeof(zc) * kkrsz * kkrsz);
  zc *tmat_n = malloc(sizeof(zc) * 4 * kkrsz * kkrsz);
  or_cgaunt(maxlmax, cgnt);
  or_ifacts(maxlmax, illp, ilp1);

  for (int j = 0; j < nrmat_ns; j++)
    for (int i = 0; i < nrmat_ns; i++) m[i + (long)ldm * j] = (i == j) ? 1.0 : 0.0;

  int nrst = 0;
  for (int ir1 = 0; ir1 < numLIZ; ir1++) {
    int kkr1 = (LIZlmax[ir1] + 1) * (LIZlmax[ir1] + 1), kkr1_ns = kkr1 * n_spin_cant;
    pack_tmat(kkrsz, kkr1, n_spin_cant, n_spin_pola, ispin,
              tmatStore + (long)ldTmatStore * LIZStoreIdx[ir1],
              (long)iie * blkSizeTmatStore, tmat_n);
    int ncst = 0;
    for (int ir2 = 0; ir2 < numLIZ; ir2++) {
      int kkr2 = (LIZlmax[ir2] + 1) * (LIZlmax[ir2] + 1), kkr2_ns = kkr2 * n_spin_cant;
      if (ir1 != ir2) {
        double rij[3];
        for (int d = 0; d < 3; d++) rij[d] = LIZPos[d + 3 * ir1] - LIZPos[d + 3 * ir2];
        or_makegij(LIZlmax[ir1], LIZlmax[ir2], maxlmax, prel, rij, cgnt, illp, ilp1, gij);
        /* M(nrst+r, ncst+c) = - sum_k tmat_n(r,k) bgij(k,c); bgij = diag_spin(gij) */
        for (int c = 0; c < kkr2_ns; c++) {
          int cs = c / kkr2, cc = c % kkr2;
          for (int r = 0; r < kkr1_ns; r++) {
            zc s = 0.0;
            for (int k = 0; k < kkr1; k++)
              s += tmat_n[r + kkr1_ns * (cs * kkr1 + k)] * gij[k + kkr1 * cc];
            m[(nrst + r) + (long)ldm * (ncst + c)] = -s;
          }
        }
      }
      ncst += kkr2_ns;
    }
    nrst += kkr1_ns;
  }
  free(cgnt); free(illp); free(gij); free(tmat_n);
  return nrmat_ns;
}

/*
 * t0 = the central site's kkrsz_ns x kkrsz_ns t-matrix
 * (linearSolvers_CPU.cpp:11-42 buildKKRSizeTMatrix).
 */
void or_t0(int maxlmax, int n_spin_cant, int n_spin_pola, int ispin,
           int storeIdx0, const zc *tmatStore, long ldTmatStore,
           long blkSizeTmatStore, int iie, zc *t0) {
  int kkrsz_ns = n_spin_cant * (maxlmax + 1) * (maxlmax + 1);
  long jsm = (n_spin_pola == n_spin_cant) ? 0 : (long)kkrsz_ns * kkrsz_ns * ispin;
  const zc *col = tmatStore + (long)ldTmatStore * storeIdx0 + (long)iie * blkSizeTmatStore + jsm;
  for (int j = 0; j < kkrsz_ns; j++)
    for (int i = 0; i < kkrsz_ns; i++) t0[i + kkrsz_ns * j] = col[i + (long)kkrsz_ns * j];
}

static double cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); }

/*
 * LU with partial pivoting, LAPACK zgetf2 semantics: pivot = first index of
 * max |re|+|im| (izamax), full-row swaps, ipiv 1-based, info = first zero
 * pivot (1-based) or 0.  LAPACK is the reference's third-party dependency
 * for this step (linearSolvers_CPU.cpp:116-118).
 */
int or_zgetrf(int m, int n, zc *a, int lda, int *ipiv) {
  int info = 0, mn = m < n ? m : n;
  for (int j = 0; j < mn; j++) {
    int p = j;
    double amax = cabs1(a[j + (long)lda * j]);
    for (int i = j + 1; i < m; i++) {
      double v = cabs1(a[i + (long)lda * j]);
      if (v > amax) { amax = v; p = i; }
    }
    ipiv[j] = p + 1;
    if (a[p + (long)lda * j] != 0.0) {
      if (p != j)
        for (int c = 0; c < n; c++) {
          zc t = a[j + (long)lda * c];
          a[j + (long)lda * c] = a[p + (long)lda * c];
          a[p + (long)lda * c] = t;
        }
      zc rp = 1.0 / a[j + (long)lda * j];
      for (int i = j + 1; i < m; i++) a[i + (long)lda * j] *= rp;
    } else if (info == 0) {
      info = j + 1;
    }
    for (int c = j + 1; c < n; c++) {
      zc u = a[j + (long)lda * c];
      if (u != 0.0) {
        zc *ac = a + (long)lda * c, *aj = a + (long)lda * j;
        for (int i = j + 1; i < m; i++) ac[i] -= aj[i] * u;
      }
    }
  }
  return info;
}

/* solve A X = B with the factors of or_zgetrf (LAPACK zgetrs 'N') */
int or_zgetrs(int n, int nrhs, const zc *a, int lda, const int *ipiv, zc *b, int ldb) {
  for (int r = 0; r < nrhs; r++) {
    zc *x = b + (long)ldb * r;
    for (int i = 0; i < n; i++) {
      int p = ipiv[i] - 1;
      if (p != i) { zc t = x[i]; x[i] = x[p]; x[p] = t; }
    }
    for (int j = 0; j < n; j++) {
      zc xj = x[j];
      if (xj != 0.0)
        for (int i = j + 1; i < n; i++) x[i] -= a[i + (long)lda * j] * xj;
    }
    for (int j = n - 1; j >= 0; j--) {
      if (x[j] != 0.0) {
        x[j] /= a[j + (long)lda * j];
        zc xj = x[j];
        for (int i = 0; i < j; i++) x[i] -= a[i + (long)lda * j] * xj;
      }
    }
  }
  return 0;
}

/* C = alpha A B + beta C, all column-major (BLAS zgemm 'N','N') */
void or_zgemm_nn(int m, int n, int k, zc alpha, const zc *a, int lda,
                 const zc *b, int ldb, zc beta, zc *c, int ldc) {
  for (int j = 0; j < n; j++) {
    zc *cj = c + (long)ldc * j;
    if (beta == 0.0)
      for (int i = 0; i < m; i++) cj[i] = 0.0;
    else if (beta != 1.0)
      for (int i = 0; i < m; i++) cj[i] *= beta;
    for (int l = 0; l < k; l++) {
      zc t = alpha * b[l + (long)ldb * j];
      if (t == 0.0) continue;
      const zc *al = a + (long)lda * l;
      for (int i = 0; i < m; i++) cj[i] += al[i] * t;
    }
  }
}

/*
 * tau00 = [M^-1 [t0;0]]_{0:k,0:k} via LU (linearSolvers_CPU.cpp:98-123).
 * M is overwritten by its factors.  Returns the zgetrf info.
 */
int or_solve_tau00_zgetrf(int nrmat_ns, int kkrsz_ns, zc *m, int ldm,
                          const zc *t0, zc *tau00) {
  int *ipiv = malloc(sizeof(int) * nrmat_ns);
  zc *tau = calloc((size_t)nrmat_ns * kkrsz_ns, sizeof(zc));
  for (int j = 0; j < kkrsz_ns; j++)
    for (int i = 0; i < kkrsz_ns; i++) tau[i + (long)nrmat_ns * j] = t0[i + kkrsz_ns * j];
  int info = or_zgetrf(nrmat_ns, nrmat_ns, m, ldm, ipiv);
  or_zgetrs(nrmat_ns, kkrsz_ns, m, ldm, ipiv, tau, nrmat_ns);
  for (int j = 0; j < kkrsz_ns; j++)
    for (int i = 0; i < kkrsz_ns; i++) tau00[i + kkrsz_ns * j] = tau[i + (long)nrmat_ns * j];
  free(ipiv);
  free(tau);
  return info;
}

/*
 * Block-LU route: zero A00, eliminate the trailing blocks from the back
 * (zblock_lu_cpp.cpp:19-86 with idcol(1)=0, i.e. k=1), delta = -A00 = B D^-1 C
 * (block_inverse.cpp:24-47), then tau00 = (1-delta)^-1 t0
 * (linearSolvers_CPU.cpp:261-331; tau_inv_postproc.f:33-57).
 * Block sizes: blk_sz[0] = kkrsz_ns, the rest split as evenly as possible
 * (linearSolvers_CPU.cpp:200-213).  Returns the first nonzero zgetrf info.
 */
int or_solve_tau00_zblocklu(int nrmat_ns, int kkrsz_ns, int nblk, zc *a, int lda,
                            const zc *t0, zc *tau00) {
  int blk[1000];
  int first_info = 0;
  if (nrmat_ns == kkrsz_ns) nblk = 1;
  blk[0] = kkrsz_ns;
  if (nblk == 1) blk[0] = nrmat_ns;
  else if (nblk == 2) blk[1] = nrmat_ns - blk[0];
  else {
    int min_sz = (nrmat_ns - blk[0]) / (nblk - 1), rem = (nrmat_ns - blk[0]) % (nblk - 1);
    int i = 1;
    for (; i <= rem; i++) blk[i] = min_sz + 1;
    for (; i < nblk; i++) blk[i] = min_sz;
  }
  int *ipvt = malloc(sizeof(int) * nrmat_ns);
  for (int j = 0; j < blk[0]; j++)
    for (int i = 0; i < blk[0]; i++) a[i + (long)lda * j] = 0.0;
  int na = 0;
  for (int i = 0; i < nblk; i++) na += blk[i];
  const int k = 1;
  if (nblk > 0) {
    int n = blk[nblk - 1], joff = na - n;
    for (int iblk = nblk - 1; iblk >= 1; iblk--) {
      int mm = n, ioff = joff;
      n = blk[iblk - 1];
      joff = joff - n;
      int info = or_zgetrf(mm, mm, a + ioff + (long)lda * ioff, lda, ipvt);
      if (info && !first_info) first_info = info;
      or_zgetrs(mm, ioff, a + ioff + (long)lda * ioff, lda, ipvt, a + ioff, lda);
      if (iblk > 1) {
        int off1 = ioff - k + 1, off2 = na - ioff;
        or_zgemm_nn(n, off1, off2, -1.0, a + joff + (long)lda * ioff, lda,
                    a + ioff + (long)lda * (k - 1), lda, 1.0,
                    a + joff + (long)lda * (k - 1), lda);
        or_zgemm_nn(joff, n, off2, -1.0, a + (long)lda * ioff, lda,
                    a + ioff + (long)lda * joff, lda, 1.0, a + (long)lda * joff, lda);
      }
    }
    int off3 = blk[0] - k + 1, off4 = na - blk[0];
    or_zgemm_nn(blk[0], off3, off4, -1.0, a + (long)lda * blk[0], lda,
                a + blk[0] + (long)lda * (k - 1), lda, 1.0, a, lda);
  }
  /* wbig = 1 - delta = 1 + A00; tau00 = wbig^-1 t0 */
  zc *wbig = malloc(sizeof(zc) * kkrsz_ns * kkrsz_ns);
  for (int j = 0; j < kkrsz_ns; j++)
    for (int i = 0; i < kkrsz_ns; i++)
      wbig[i + kkrsz_ns * j] = (i == j ? 1.0 : 0.0) + a[i + (long)lda * j];
  int info = or_zgetrf(kkrsz_ns, kkrsz_ns, wbig, kkrsz_ns, ipvt);
  if (info && !first_info) first_info = info;
  memcpy(tau00, t0, sizeof(zc) * kkrsz_ns * kkrsz_ns);
  or_zgetrs(kkrsz_ns, kkrsz_ns, wbig, kkrsz_ns, ipvt, tau00, kkrsz_ns);
  free(wbig);
  free(ipvt);
  return first_info;
}

/* tau00 - t0 (tau00Postprocess.cpp:52-83) */
void or_tau00_minus_t(int kkrsz_ns, const zc *tau00, const zc *t0, zc *out) {
  for (int i = 0; i < kkrsz_ns * kkrsz_ns; i++) out[i] = tau00[i] - t0[i];
}

/* spin rotation u, u^dagger for a moment direction (Misc/spin_trafo.f) */
void or_spin_trafo(const double *evec, zc *u, zc *ud) {
  u[0] = sqrt(0.5 * (1.0 + evec[2]));
  if (cabs(u[0]) < 1.0e-8) {
    u[0] = 0.0;
    u[1] = 1.0;
    u[2] = 1.0;
  } else {
    u[1] = -0.5 * (evec[0] + I * evec[1]) / u[0];
    u[2] = 0.5 * (evec[0] - I * evec[1]) / u[0];
  }
  u[3] = u[0];
  ud[0] = conj(u[0]);
  ud[1] = conj(u[2]);
  ud[2] = conj(u[1]);
  ud[3] = conj(u[3]);
}

/*
 * Rotate a 2nr x 2nc spin matrix to the local frame: the four nr x nc
 * blocks ml(:,:,k), k = (s,s') in column-major order (1,1),(2,1),(1,2),(2,2),
 * ml_{ss'} = sum_{ab} U(s,a) mg_{ab} UD(b,s') with U = [[u1,u3],[u2,u4]]
 * (MultipleScattering/trgtol.f:3-110).
 */
void or_trgtol(int nr, int nc, const zc *u, const zc *ud, const zc *mg, zc *ml) {
  /* U(s,a) = u[s + 2a], UD(b,s') = ud[b + 2 s'] */
  for (int sp = 0; sp < 2; sp++)
    for (int s = 0; s < 2; s++) {
      zc *blk = ml + (long)nr * nc * (s + 2 * sp);
      for (int c = 0; c < nc; c++)
        for (int r = 0; r < nr; r++) {
          zc acc = 0.0;
          for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++)
              acc += u[s + 2 * a] * ud[b + 2 * sp] *
                     mg[(a * nr + r) + (long)(2 * nr) * (b * nc + c)];
          blk[r + (long)nr * c] = acc;
        }
    }
}

/*
 * Semicircle Gauss-Legendre contour from ebot to etop
 * (Misc/congauss_c.f:3-124): E_k = mid + r e^{i pi (1-x_k)/2},
 * w_k = -i (pi/2) (E_k - mid) wgs_k.  The reference appends one extra point
 * etop + i eibot with weight 0; it carries no weight and is not produced here.
 */
void or_congauss(double ebot, double etop, int ngauss, zc *egrd, zc *dele) {
  const double pi = 2.0 * asin(1.0); /* energyContourIntegration.cpp:89 */
  double *x = malloc(sizeof(double) * ngauss), *w = malloc(sizeof(double) * ngauss);
  or_gauss_legendre(ngauss, -1.0, 1.0, x, w);
  double erad = 0.5 * (etop - ebot);
  zc ihalfpi = I * 0.5 * pi;
  for (int ie = 0; ie < ngauss; ie++) {
    zc zz = erad * cexp(ihalfpi * (1.0 - x[ie]));
    egrd[ie] = 0.5 * (etop + ebot) + zz;
    dele[ie] = -ihalfpi * zz * w[ie];
  }
  free(x);
  free(w);
}

/* ---- relativistic KKR blocks (relativity = full) ------------------------ */

/*
 * Clebsch-Gordan tables of the (l m s) -> (kappa mu) change of basis
 * (Misc/clebsch.f:1-60): for every l <= 12 and m, the kappa-mu row
 * ir = 2 kappa^2 + kappa + m (+1 for the spin-up component), 1-based, gets
 *   j = l - 1/2 (kappa = l, l > 0):  u1 = sqrt((l+m)/(2l+1)),  u2 = -sqrt((l-m)/(2l+1))
 *   j = l + 1/2 (kappa = -l-1):      u1 = sqrt((l-m+1)/(2l+1)), u2 = sqrt((l+m+1)/(2l+1))
 * with ind1/ind2 = the 1-based lm index l^2 + l + m + 1; an index whose
 * coefficient is below 1e-15 is set to 1.  Arrays of 500 as the common block
 * /cgc/ (Misc/cgc.h); element ir is at [ir - 1].
 */
void or_clebsch(double *u1, double *u2, int *ind1, int *ind2) {
  for (int i = 0; i < 500; i++) {
    u1[i] = 0.0;
    u2[i] = 0.0;
    ind1[i] = 1;
    ind2[i] = 1;
  }
  int inr = 0;
  for (int l = 0; l <= 12; l++) {
    const double twolp1 = 2.0 * l + 1.0;
    for (int m = -l; m <= l; m++) {
      inr++;
      if (l > 0) {
        const int kap = l;
        int ir = 2 * kap * kap + kap + m;
        u1[ir - 1] = sqrt((double)(l + m) / twolp1);
        ind1[ir - 1] = inr;
        ir++;
        u2[ir - 1] = -sqrt((double)(l - m) / twolp1);
        ind2[ir - 1] = inr;
      }
      const int kap = -l - 1;
      int ir = 2 * kap * kap + kap + m;
      u1[ir - 1] = sqrt((double)(l - m + 1) / twolp1);
      ind1[ir - 1] = inr;
      ir++;
      u2[ir - 1] = sqrt((double)(l + m + 1) / twolp1);
      ind2[ir - 1] = inr;
    }
  }
  for (int ir = 0; ir < inr; ir++) {
    if (fabs(u1[ir]) < 1e-15) ind1[ir] = 1;
    if (fabs(u2[ir]) < 1e-15) ind2[ir] = 1;
  }
}

/*
 * kappa-mu form of a non-relativistic kkr x kkr block a
 * (MultipleScattering/relmtrx.f:1-45):
 *   b(i, j) = u1(i) a(ind1(i), ind1(j)) u1(j) + u2(i) a(ind2(i), ind2(j)) u2(j),
 * i, j = 1 .. 2 kkr (b is 2kkr x 2kkr, column-major).
 */
void or_relmtrx(const zc *a, int kkr, zc *b) {
  double u1[500], u2[500];
  int ind1[500], ind2[500];
  or_clebsch(u1, u2, ind1, ind2);
  const int n = 2 * kkr;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      const zc t1 = u1[i] * a[(ind1[i] - 1) + (long)kkr * (ind1[j] - 1)] * u1[j];
      const zc t2 = u2[i] * a[(ind2[i] - 1) + (long)kkr * (ind2[j] - 1)] * u2[j];
      b[i + (long)n * j] = t1 + t2;
    }
}

/*
 * Relativistic KKR matrix, F77 route (calculateTauMatrix.cpp:216-242 with
 * nrel_rel = 1; setgij.f:39-55): every LIZ site at lmax = maxlmax,
 * bgij = (prel^2 / E) relmtrx(gij), M(site i, site j) = -t_i bgij with t_i
 * the full 2 kkrsz x 2 kkrsz kappa-mu t-matrix of the store.  Returns
 * nrmat_ns (= 2 nrmat), or -1 if a site has another lmax.
 */
int or_build_kkr_rel(int maxlmax, int numLIZ, const double *LIZPos, const int *LIZlmax,
                     const int *LIZStoreIdx, const zc *tmatStore, long ldTmatStore,
                     long blkSizeTmatStore, int iie, zc prel, zc energy, zc *m, int ldm) {
  for (int s = 0; s < numLIZ; s++)
    if (LIZlmax[s] != maxlmax) return -1;
  const int kkr = (maxlmax + 1) * (maxlmax + 1), kk2 = 2 * kkr;
  const int nrmat_ns = kk2 * numLIZ;
  double *cgnt = malloc(sizeof(double) * (maxlmax + 1) * kkr * kkr);
  zc *illp = malloc(sizeof(zc) * kkr * kkr);
  zc ilp1[64];
  zc *gij = malloc(sizeof(zc) * kkr * kkr);
  zc *bgij = malloc(sizeof(zc) * kk2 * kk2);
  zc *tmat_n = malloc(sizeof(zc) * kk2 * kk2);
  or_cgaunt(maxlmax, cgnt);
  or_ifacts(maxlmax, illp, ilp1);
  const zc fac = prel * prel / energy;
  for (int j = 0; j < nrmat_ns; j++)
    for (int i = 0; i < nrmat_ns; i++) m[i + (long)ldm * j] = (i == j) ? 1.0 : 0.0;
  for (int ir1 = 0; ir1 < numLIZ; ir1++) {
    pack_tmat(kkr, kkr, 2, 2, 0, tmatStore + (long)ldTmatStore * LIZStoreIdx[ir1],
              (long)iie * blkSizeTmatStore, tmat_n);
    for (int ir2 = 0; ir2 < numLIZ; ir2++) {
      if (ir1 == ir2) continue;
      double rij[3];
      for (int d = 0; d < 3; d++) rij[d] = LIZPos[d + 3 * ir1] - LIZPos[d + 3 * ir2];
      or_makegij(maxlmax, maxlmax, maxlmax, prel, rij, cgnt, illp, ilp1, gij);
      or_relmtrx(gij, kkr, bgij);
      for (int e = 0; e < kk2 * kk2; e++) bgij[e] = fac * bgij[e];
      for (int c = 0; c < kk2; c++)
        for (int r = 0; r < kk2; r++) {
          zc s = 0.0;
          for (int k = 0; k < kk2; k++) s += tmat_n[r + kk2 * k] * bgij[k + kk2 * c];
          m[(kk2 * ir1 + r) + (long)ldm * (kk2 * ir2 + c)] = -s;
        }
    }
  }
  free(cgnt); free(illp); free(gij); free(bgij); free(tmat_n);
  return nrmat_ns;
}

/*
 * b <- u b ust for n x n matrices (SingleSite/matops.f tripmt): the local
 * frame of the relativistic tau00, tau00_l = dmatp tau00 dmat
 * (tau00Postprocess.cpp:185-192; tau_inv_postproc.f:79-130).
 */
void or_tripmt(const zc *u, zc *b, const zc *ust, int n) {
  zc *c = malloc(sizeof(zc) * n * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      zc x = 0.0;
      for (int k = 0; k < n; k++) x += b[k + (long)n * j] * u[i + (long)n * k];
      c[i + (long)n * j] = x;
    }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      zc x = 0.0;
      for (int k = 0; k < n; k++) x += ust[k + (long)n * j] * c[i + (long)n * k];
      b[i + (long)n * j] = x;
    }
  free(c);
}
