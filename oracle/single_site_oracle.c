/*
 * single_site_oracle.c -- CPU restatement of the reference's single-site
 * scattering solve for non-relativistic and scalar-relativistic potentials:
 * t_l = 1/matom_l and the radial functions zlr, jlr.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in
 * lsms_amd/csrc/single_site.hip.  Only tests/, smoke() and bench.py's CPU
 * baseline load it.
 *
 * Follows, step for step (the reference's semrel uses solver 2):
 *   single_scatterer_nonrel   SingleSite/single_scatterer_nonrel.f:4-122
 *   semrel (solver = 2)       SingleSite/semrel_c.f:3-228
 *   fitpot                    Misc/fitpot.f:2-88
 *   rwave                     Misc/rwave.f:1-242
 *   initwave                  Misc/initwave.f:1-109
 *   dfv / dfv_m (nv = 4)      Misc/dfv.f:1-71, Misc/dfv_new.f:1-71
 *   Bulirsch-Stoer step       Misc/bulirschStoerIntegrator.cpp:20-141,
 *                             Misc/bulirschStoerIntegrator.hpp:40-88 (extrapolation)
 *   cinterp / fit             Misc/cinterp.f:1-50, Misc/fit.f:23-131
 *   zsphbesjh / zsphbesj      Misc/zsphbesjh.f:1-45, Misc/zsphbesj.f:1-55
 * Indices are 0-based here; the reference's grid point j is index j-1.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lsms_oracle.h"

/* ---- fitpot.f:21-86: rv on [r_j, r_j+1] as
 *   (c0 (r-r_j) + c1)/(1 + c2 (r-r_j)^2) + c3,
 * or linear (c2 = c3 = 0) where the rational fit degenerates.  `nfit` is the
 * routine's own `jmt` argument (semrel passes iend-1); points nfit-1..iend-1
 * get zero coefficients. */
static void fitpot4(const double *r, const double *rv, double *coef, int nfit, int iend)
{
    for (int j = 0; j < nfit - 1; j++) {
        double r0 = r[j], dr = r[j + 1] - r0;
        double drvdr = (rv[j + 1] - rv[j]) / dr;
        double *c = coef + 4 * j;
        /* outside points: j-1 and j+2; 2 and 3 on the first interval; j-2 and
         * j-1 on the last (fitpot.f:40-46) */
        int i1 = (j == 0) ? 2 : j - 1;
        if (j == nfit - 2) i1 -= 1;
        double h1 = r[i1] - r0, h2 = r[i1] - r[j + 1];
        double e11 = h2, e12 = -h1 * (rv[i1] - rv[j + 1]);
        double rhs1 = (rv[i1] - rv[j]) / h1 - drvdr;
        i1 += 1;
        if (i1 == j) i1 += 2;
        h1 = r[i1] - r0;
        h2 = r[i1] - r[j + 1];
        double e21 = h2, e22 = -h1 * (rv[i1] - rv[j + 1]);
        double rhs2 = (rv[i1] - rv[j]) / h1 - drvdr;
        double det = e11 * e22 - e21 * e12;
        int linear = 1;
        if (fabs(det) > 1e-16) {
            c[1] = -(e12 * rhs2 - e22 * rhs1) / det;
            c[2] = (e11 * rhs2 - e21 * rhs1) / det;
            double t = c[2] * dr * dr;
            if (t > -1.0 && fabs(t) > 1e-14) {
                c[0] = drvdr - c[1] * dr;
                c[3] = rv[j + 1] + c[1] / c[2];
                c[1] = rv[j] - c[3];
                linear = 0;
            }
        }
        if (linear) {
            c[0] = drvdr;
            c[1] = rv[j];
            c[2] = 0.0;
            c[3] = 0.0;
        }
    }
    for (int j = (nfit > 1 ? nfit - 1 : 0); j < iend; j++)
        for (int i = 0; i < 4; i++) coef[4 * j + i] = 0.0;
}

/* ---- fit.f:35-131: four-point rational interpolation coefficients around
 * point i (0-based), f(i) - c0 + c1 (r-r_i) + ((c2-c1)(r-r_i) + c0)/(1 + c3 (r-r_i)(r_i+1 - r)). */
static int fit4(const double *r, const double *f, int nr, int i, double c[4])
{
    int ip1 = i + 1;
    if (i < 0 || i >= nr) return -1;
    if (ip1 >= nr) { /* wrap around on a ring of points (fit.f:53-62) */
        if (r[0] > r[nr - 1])
            ip1 = 0;
        else
            return -1;
    }
    double dr = r[ip1] - r[i];
    if (dr == 0.0) return -1;
    c[2] = (f[ip1] - f[i]) / dr;
    int i1 = (i > 0) ? i - 1 : (nr < 4 ? nr : 4) - 1;
    int i2 = (ip1 < nr - 1) ? ip1 + 1 : (nr - 3 > 1 ? nr - 3 : 1) - 1;
    if (i1 != i2) {
        c[1] = (f[i2] - f[ip1]) / (r[i2] - r[ip1]) - (f[i1] - f[i]) / (r[i1] - r[i]);
        if (c[1] * (f[i2] - f[i1]) * (r[i2] - r[i1]) > 0.0) c[1] = -c[1];
        double h1 = r[i1] - r[i];
        double d0 = f[i1] - f[i];
        double a1 = d0 - c[2] * h1, a2 = d0 - c[1] * h1;
        h1 = h1 * (r[ip1] - r[i1]);
        double e12 = -h1 * a2;
        double h2 = r[i2] - r[i];
        d0 = f[i2] - f[i];
        double a3 = d0 - c[2] * h2, a4 = d0 - c[1] * h2;
        h2 = h2 * (r[ip1] - r[i2]);
        double e22 = -h2 * a4;
        double gj = (a4 - a2) * h1 * h2;
        if (gj == 0.0) {
            c[3] = 0.0;
        } else {
            c[0] = (a1 * e22 - a3 * e12) / gj;
            c[3] = (a1 * h2 - a3 * h1) / gj;
        }
        gj = c[3] * dr * dr;
        if (gj > -4.0 && fabs(gj) > 1e-14 && gj < 1e14) {
            c[0] = c[0] / c[3];
        } else {
            c[0] = f[i];
            c[3] = 0.0;
        }
    } else {
        double fj = 0.5 * (f[i] + f[ip1]);
        c[0] = f[i] - fj;
        c[1] = 0.0;
        double rj = r[i] + 0.5 * dr;
        if (fabs(r[i1] - rj) > fabs(r[i2] - rj)) i1 = i2;
        double h1 = r[i1] - r[i];
        double a1 = f[i1] - f[i] - c[2] * h1;
        h1 = h1 * (r[ip1] - r[i1]);
        if (f[i1] != fj) {
            c[3] = -a1 / ((f[i1] - fj) * h1);
            double gj = c[3] * dr * dr;
            if (gj <= -4.0 || fabs(gj) <= 1e-14 || gj >= 1e14) {
                c[0] = f[i];
                c[3] = 0.0;
            }
        } else {
            c[0] = f[i];
            c[3] = 0.0;
        }
    }
    return 0;
}

/* ---- cinterp.f:1-50: complex f on r[0..nr) interpolated at rs (value only). */
static zc cinterp_value(const double *r, const zc *f, int nr, double rs, double *work)
{
    int ip1 = 1;
    for (int i = 1; i <= nr - 2; i++)
        if (rs > r[i]) ip1 = i + 1;
    int i = ip1 - 1;
    double part[2];
    for (int p = 0; p < 2; p++) {
        for (int k = 0; k < nr; k++) work[k] = p ? cimag(f[k]) : creal(f[k]);
        double c[4] = {0.0, 0.0, 0.0, 0.0};
        fit4(r, work, nr, i, c);
        double h1 = rs - r[i], h2 = r[ip1] - rs;
        double den = 1.0 + c[3] * h1 * h2;
        double ps = ((c[2] - c[1]) * h1 + c[0]) / den;
        part[p] = ps + work[i] - c[0] + c[1] * h1;
    }
    return part[0] + I * part[1];
}

/* ---- zsphbesj.f / zsphbesjh.f: j_l(x) and the retarded Hankel function
 * h_l(x) = hm[l] e^{ix}/x, l = 0..lmax (lmax >= 1). */
static void sph_bessel_jh(int lmax, zc x, zc *bj, zc *hm, zc *eiz)
{
    zc bs[16];
    zc as = csin(x), ac = ccos(x), x1 = 1.0 / x;
    bj[0] = as * x1;
    int k = ((int)(0.75 * (fabs(creal(x)) + fabs(cimag(x)))) + 2) >> 2;
    bs[0] = as;
    int done = 0;
    if (k >= 1) {
        if (k > lmax) k = lmax;
        bs[1] = bj[0] - ac;
        bj[1] = bs[1] * x1;
        for (int l = 2; l <= k; l++) {
            bs[l] = (double)(2 * l - 1) * bj[l - 1] - bs[l - 2];
            bj[l] = bs[l] * x1;
        }
        if (k == lmax) done = 1;
    }
    if (!done) {
        zc a = bs[k];
        int nm = lmax << 4;
        zc aj = (double)(2 * nm + 3);
        zc x2 = x * x;
        for (int l = nm; l >= lmax + 2; l--) aj = (double)(2 * l + 1) - x2 / aj;
        bs[lmax] = (double)(2 * lmax + 3) * aj * x1 - x;
        bj[lmax] = bs[lmax] * x1;
        bs[lmax - 1] = (double)(2 * lmax + 1) * bj[lmax] - aj;
        for (int l = lmax - 1; l >= k + 1; l--) {
            bj[l] = bs[l] * x1;
            bs[l - 1] = (double)(2 * l + 1) * bj[l] - bs[l + 1];
        }
        aj = a / bs[k];
        for (int l = k + 1; l <= lmax; l++) bj[l] = aj * bj[l];
    }
    *eiz = cexp(I * x);
    hm[0] = -I;
    hm[1] = -(1.0 - cimag(x1)) - I * creal(x1);
    for (int l = 2; l <= lmax; l++) hm[l] = (double)(2 * l - 1) * hm[l - 1] * x1 - hm[l - 2];
}

/* ---- dfv_new.f (nv = 4): derivatives of (Re g, Im g, Re f, Im f) for the
 * scalar-relativistic radial equation, potential r v(r) from the fit at rn. */
typedef struct {
    zc e;
    double b, allp1;
} rhs_par;

static void dfv4(double r, const double *y, double *dy, double rn, const double *pc,
                 const rhs_par *p)
{
    double s1or = 1.0 / r, s1or2 = s1or * s1or;
    double r1 = r - rn;
    double v = (pc[0] * r1 + pc[1]) / (1.0 + pc[2] * r1 * r1) + pc[3];
    double s2vmer = v * s1or - creal(p->e);
    double frelr = 1.0 - p->b * s2vmer;
    double s2vmei = -cimag(p->e);
    double freli = -p->b * s2vmei;
    double den = p->allp1 / (frelr * frelr + freli * freli) * s1or2;
    double facr = den * frelr + s2vmer;
    double faci = -den * freli + s2vmei;
    dy[0] = frelr * y[2] - freli * y[3] + s1or * y[0];
    dy[1] = freli * y[2] + frelr * y[3] + s1or * y[1];
    dy[2] = facr * y[0] - faci * y[1] - y[2] * s1or;
    dy[3] = faci * y[0] + facr * y[1] - y[3] * s1or;
}

/* ---- modified midpoint (bulirschStoerIntegrator.cpp:26-56) */
static void midpoint4(double x0, double x1, const double *y0, double *y1, int steps, double rn,
                      const double *pc, const rhs_par *p)
{
    double y2[4], dy[4];
    double h = (x1 - x0) / (double)steps, h2 = 2.0 * h;
    double x = x0;
    dfv4(x, y0, dy, rn, pc, p);
    for (int i = 0; i < 4; i++) {
        y1[i] = y0[i];
        y2[i] = y0[i] + h * dy[i];
    }
    for (int j = 1; j < steps; j++) {
        x += h;
        dfv4(x, y2, dy, rn, pc, p);
        for (int i = 0; i < 4; i++) {
            double t = y1[i] + h2 * dy[i];
            y1[i] = y2[i];
            y2[i] = t;
        }
    }
    x += h;
    dfv4(x, y2, dy, rn, pc, p);
    for (int i = 0; i < 4; i++) y1[i] = 0.5 * (y1[i] + y2[i] + h * dy[i]);
}

/* ---- adaptive Bulirsch-Stoer over [x0, x1] with polynomial extrapolation in
 * h^2 (bulirschStoerIntegrator.cpp:58-122; tableau ExtrapolatorT0,
 * bulirschStoerIntegrator.hpp:40-88).  Returns 0, or 1 if the step underflowed. */
#define BS_NSTEPS 11
static int bs_integrate4(double x0, double x1, const double *y0, double *yout, double rn,
                         const double *pc, const rhs_par *p)
{
    static const int seq[] = {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128};
    const double eps = 1.0e-12;
    double T[(BS_NSTEPS + 2) * (BS_NSTEPS + 3) / 2][4];
    double hh[BS_NSTEPS + 2];
    int knum = 0;
    double y[4], y1[4];
    double xeps = 1.0e-12 * fabs(x0);
    double x = x0, h0 = x1 - x0, h = h0;
    int rc = 0;
    memcpy(y, y0, sizeof y);
#define TIX(i, k) ((i) * ((i) + 1) / 2 + (k))
    do {
        int step = 0;
        double err = HUGE_VAL;
        knum = 0;
        while (step < BS_NSTEPS && err > eps) {
            midpoint4(x, x + h, y, y1, seq[step], rn, pc, p);
            int i = knum;
            hh[i] = h / (double)seq[step];
            knum++;
            for (int d = 0; d < 4; d++) {
                T[TIX(i, 0)][d] = y1[d];
                for (int k = 1; k < knum; k++) {
                    double rr = hh[i - k] / hh[i];
                    double den = 1.0 / (rr * rr - 1.0);
                    T[TIX(i, k)][d] = T[TIX(i, k - 1)][d] +
                                      (T[TIX(i, k - 1)][d] - T[TIX(i - 1, k - 1)][d]) * den;
                }
            }
            double e = 0.0, mag = 0.0;
            if (knum < 2) {
                e = 4.0;
            } else {
                for (int d = 0; d < 4; d++)
                    e += fabs(T[TIX(knum - 1, knum - 1)][d] - T[TIX(knum - 2, knum - 2)][d]);
            }
            for (int d = 0; d < 4; d++) mag += fabs(T[TIX(knum - 1, knum - 1)][d]);
            if (mag == 0.0) mag = 1.0e-15;
            err = e / mag;
            step++;
        }
        if (err < eps) {
            x = x + h;
            if (step < 7)
                h *= 1.25;
            else
                h *= (double)seq[7] / (double)seq[step];
            double lim = fabs(x1 - x);
            h = copysign(fabs(h) < lim ? fabs(h) : lim, h0);
            for (int d = 0; d < 4; d++) y[d] = T[TIX(knum - 1, knum - 1)][d];
        } else {
            h = 0.25 * h;
            if (h + x == x) {
                rc = 1;
                break;
            }
        }
    } while (fabs(x - x1) > xeps);
    for (int d = 0; d < 4; d++) yout[d] = T[TIX(knum - 1, knum - 1)][d];
#undef TIX
    return rc;
}

/* one grid interval as the reference's bulirsch_stoer_integrator_ does it:
 * integrate, then the derivative at the end point (cpp:124-141) */
static int bs_interval(double xfrom, double xto, double *y, double *dy, double rn,
                       const double *pc, const rhs_par *p)
{
    double y1[4];
    int rc = bs_integrate4(xfrom, xto, y, y1, rn, pc, p);
    memcpy(y, y1, sizeof y1);
    dfv4(xto, y, dy, rn, pc, p);
    return rc;
}

/* ---- rwave.f:1-242 (irregflg = true, potshift = 0, emach = 1e-14, eunit = 1):
 * regular solution rr/drr and irregular ri/dri on rad[0..kmax), matom = 1/t_l. */
int or_rwave(int srel, int l, zc energy, const double *coef, const double *rad, int jmt,
             int kmax, double r_sph, zc *rr, zc *drr, zc *ri, zc *dri, zc *matom)
{
    const double c = 137.0359895, c2 = c * c; /* initwave.f:29-30 */
    const double escale = 0.5;
    zc ka = csqrt(energy + I * 1.0e-14);
    double al = (double)l;
    int istart = 0;
    if (rad[0] == 0.0) {
        rr[0] = 0.0;
        drr[0] = (l == 0) ? 1.0 : 0.0;
        istart = 1;
    }
    /* initwave.f:45-97 (core = false) */
    rhs_par p;
    p.e = energy;
    double y[4], dy[4];
    {
        const double *pot = coef; /* point 1's fit */
        double z0 = pot[1] + pot[3];
        z0 = (double)(int)(fabs(z0) + 0.1);
        double z = escale * z0;
        int rel = srel && !(z < 0.5);
        p.allp1 = fabs((double)(l * (l + 1)));
        double x = rad[istart];
        zc g1, g2;
        if (rel) {
            p.b = 0.25 / c2;
            g1 = 1.0;
            g2 = 2.0 * c2 * (sqrt(p.allp1 + 1.0 - z * z / c2) - 1.0) / z;
        } else {
            p.b = 0.0;
            zc v0 = escale * (energy - pot[0]);
            double v1 = 2.0 * escale * pot[1] * pot[2];
            v0 = v0 - 0.5 * v1 * (rad[0] + rad[1]);
            double alp1 = al + 1.0;
            double a1 = -z / alp1;
            zc a2 = (-z * a1 - v0) / (2.0 * al + 3.0);
            zc a3 = -(-a1 * (z * z - (3.0 * al + 4.0) * v0) + (2.0 * al + 3.0) * v1) /
                    (3.0 * (al + 2.0) * (2.0 * al + 3.0));
            double rl = pow(x, al);
            g1 = (x * rl) * (1.0 + x * (a1 + x * (a2 + x * a3)));
            g2 = rl * (al + x * (alp1 * a1 + x * ((al + 2.0) * a2 + x * (al + 3.0) * a3)));
        }
        y[0] = creal(g1);
        y[1] = cimag(g1);
        y[2] = creal(g2);
        y[3] = cimag(g2);
    }
    int rc = 0;
    dfv4(rad[istart], y, dy, rad[istart], coef + 4 * istart, &p);
    rr[istart] = y[0] + I * y[1];
    drr[istart] = dy[0] + I * dy[1];
    for (int j = istart + 1; j < jmt; j++) {
        rc |= bs_interval(rad[j - 1], rad[j], y, dy, rad[j - 1], coef + 4 * (j - 1), &p);
        rr[j] = y[0] + I * y[1];
        drr[j] = dy[0] + I * dy[1];
    }
    /* match to j_l + tatoml h_l at r_sph (rwave.f:123-171) */
    double *work = (double *)malloc(sizeof(double) * (size_t)jmt);
    zc drrsph = cinterp_value(rad, drr, jmt, r_sph, work);
    zc rrsph = cinterp_value(rad, rr, jmt, r_sph, work);
    free(work);
    zc dfi = drrsph / rrsph, fi = rrsph / r_sph;
    zc kar = ka * r_sph;
    int lb = l > 1 ? l : 1;
    zc bjv[16], hm[16], eiz;
    sph_bessel_jh(lb, kar, bjv, hm, &eiz);
    zc bj = bjv[l], bh1 = hm[l] * eiz / ka, dbj, dbh1;
    if (l == 0) {
        dbj = -ka * (bjv[1] - bjv[0] / kar);
        dbh1 = -(hm[1] - hm[0] / kar) * eiz;
    } else {
        zc cd = al / kar;
        dbj = ka * (bjv[l - 1] - bjv[l] * cd);
        dbh1 = (hm[l - 1] - hm[l] * cd) * eiz;
    }
    zc cdum = -(bj * dfi - dbj) / (bh1 * dfi - dbh1);
    zc cnorm = (bj + cdum * bh1) / fi;
    zc tatoml = r_sph * cdum;
    *matom = -I * ka / tatoml;
    for (int j = 0; j < jmt; j++) {
        rr[j] = cnorm * rr[j];
        drr[j] = cnorm * drr[j];
    }
    /* irregular solution, inwards from r_sph (rwave.f:181-223) */
    kar = ka * r_sph;
    sph_bessel_jh(lb, kar, bjv, hm, &eiz);
    bh1 = hm[l] * eiz / ka;
    if (l == 0) {
        dbh1 = -hm[1] * eiz;
    } else {
        zc cd = (double)(l + 1) / kar;
        dbh1 = (hm[l - 1] - hm[l] * cd) * eiz;
    }
    y[0] = creal(bh1);
    y[1] = cimag(bh1);
    y[2] = creal(dbh1);
    y[3] = cimag(dbh1);
    dfv4(r_sph, y, dy, r_sph, coef + 4 * (jmt - 1), &p);
    ri[jmt - 1] = y[0] + I * y[1];
    dri[jmt - 1] = dy[0] + I * dy[1];
    if (rad[0] == 0.0) {
        ri[0] = 0.0;
        dri[0] = 0.0;
    }
    for (int j = jmt - 2; j >= istart; j--) {
        double rfrom = (j + 1 == jmt - 1) ? r_sph : rad[j + 1];
        rc |= bs_interval(rfrom, rad[j], y, dy, rad[j], coef + 4 * j, &p);
        ri[j] = y[0] + I * y[1];
        dri[j] = dy[0] + I * dy[1];
    }
    /* outside the sphere: Bessel and Hankel forms (rwave.f:225-240) */
    for (int i = jmt - 1; i < kmax; i++) {
        kar = rad[i] * ka;
        sph_bessel_jh(lb, kar, bjv, hm, &eiz);
        bj = bjv[l];
        bh1 = hm[l] * eiz / ka;
        if (l == 0) {
            dbj = -ka * (bjv[1] - bjv[0] / kar);
            dbh1 = -(hm[1] - hm[0] / kar) * eiz;
        } else {
            zc cd = al / kar;
            dbj = ka * (bjv[l - 1] - bjv[l] * cd);
            dbh1 = (hm[l - 1] - hm[l] * cd) * eiz;
        }
        rr[i] = bj * rad[i] + bh1 * tatoml;
        ri[i] = bh1;
        drr[i] = dbj * rad[i] + dbh1 * tatoml;
        dri[i] = dbh1;
    }
    return rc;
}

/* ---- single_scatterer_nonrel.f:4-122 through semrel_c.f (solver 2):
 * matom(0..lmax), tmat (kkrsz x kkrsz, diagonal 1/matom_l), and
 * zlr(j,l) = g_l(r_j) matom_l / r_j, jlr(j,l) = i prel gpl_l(r_j) / (r_j matom_l)
 * for j < jws (column-major iprpts x (lmax+1)).  tmat/zlr/jlr may be NULL.
 * Returns 0; -1 on bad sizes (iend > iprpts, lmax < 1); 1 if an interval's
 * Bulirsch-Stoer step underflowed (the reference prints and continues). */
int or_single_scatterer_nonrel(int nrelv, int lmax, zc energy, zc prel, const double *vr,
                               const double *r, int jmt, int jws, double r_sph, int iprpts,
                               zc *matom, zc *tmat, zc *zlr, zc *jlr)
{
    int iend = jmt + 11 > jws ? jmt + 11 : jws;
    if (iend > iprpts || lmax < 1 || lmax > 14 || jmt < 5) return -1;
    double *coef = (double *)calloc((size_t)4 * iend, sizeof(double));
    zc *g = (zc *)malloc(sizeof(zc) * 4 * (size_t)iend);
    zc *f = g + iend, *gpl = g + 2 * iend, *fpl = g + 3 * iend;
    fitpot4(r, vr, coef, iend - 1, iend);
    int rc = 0;
    for (int l = 0; l <= lmax; l++) {
        rc |= or_rwave(nrelv == 0, l, energy, coef, r, jmt, iend, r_sph, g, f, gpl, fpl, &matom[l]);
        for (int j = 0; j < jws; j++) {
            if (zlr) zlr[j + (size_t)l * iprpts] = g[j] * matom[l] / r[j];
            if (jlr) jlr[j + (size_t)l * iprpts] = I * prel * gpl[j] / (r[j] * matom[l]);
        }
    }
    if (tmat) {
        int kkrsz = (lmax + 1) * (lmax + 1);
        for (int i = 0; i < kkrsz * kkrsz; i++) tmat[i] = 0.0;
        int lm = 0;
        for (int l = 0; l <= lmax; l++)
            for (int m = -l; m <= l; m++, lm++) tmat[lm + (size_t)lm * kkrsz] = 1.0 / matom[l];
    }
    free(coef);
    free(g);
    return rc;
}

/* fitpot alone, for the tests of the potential fit (coef: 4 x iend). */
void or_fitpot(const double *r, const double *rv, double *coef, int nfit, int iend)
{
    fitpot4(r, rv, coef, nfit, iend);
}

/* spherical Bessel / Hankel pair, for the tests: bj, hm of length lmax+1. */
void or_sph_bessel_jh(int lmax, zc x, zc *bj, zc *hm, zc *eiz)
{
    sph_bessel_jh(lmax, x, bj, hm, eiz);
}

/* ---- cnewint.f: running integral g(k) = int_{x_1}^{x_k} f(x) x^ip0 dx on an
 * arbitrary increasing mesh x (here x = sqrt(r)), from the rational
 * four-point fit of f in s = sqrt(x) on a ring of four points, its linear part
 * integrated exactly and the rest by Simpson's rule on six sub-intervals. */
void or_cnewint(int nr, const double *x, const zc *f, zc *g, int ip0)
{
    const int nj = 6;
    const double dnj = 1.0 / nj, wt0 = 2.0 / (3.0 * nj), wt1 = 3.0 * wt0;
    const int ip = 2 * ip0 + 1;
    const double pip1 = 1.0 / (double)(ip + 1), pip2 = 1.0 / (double)(ip + 2);
    double rs[4], fr[4], fi[4], cr[4], ci[4];
    g[0] = 0.0;
    int i1 = 0, i2 = 1, i3 = 2, i4 = 3, i0 = 0; /* ring slots */
    for (int k = 0; k < 4; k++) {
        rs[k] = sqrt(x[k]);
        fr[k] = creal(f[k]);
        fi[k] = cimag(f[k]);
    }
    int ip1 = 1, ip2 = 4; /* 1-based point counters */
    (void)i1;
    (void)i2;
    (void)i3;
    for (int i = 1; i <= nr - 1; i++) {
        ip1 += 1;
        if (i > 2 && ip2 <= nr) {
            rs[i4] = sqrt(x[ip2 - 1]);
            fr[i4] = creal(f[ip2 - 1]);
            fi[i4] = cimag(f[ip2 - 1]);
        }
        fit4(rs, fr, 4, i0, cr);
        fit4(rs, fi, 4, i0, ci);
        const double rip1 = rs[(i0 + 1) % 4];
        double gj = 1.0, fj = rip1;
        for (int j = 1; j <= ip; j++) {
            gj = gj * rs[i0] + fj;
            fj = fj * rip1;
        }
        fj = (gj * rs[i0] + fj) * pip2;
        gj = gj * pip1;
        double gr = (fr[i0] - cr[2] * rs[i0]) * gj + cr[2] * fj;
        double gi = (fi[i0] - ci[2] * rs[i0]) * gj + ci[2] * fj;
        const double dr = rip1 - rs[i0];
        const double h1 = dnj * dr;
        double h2 = 0.0, wt = wt0;
        for (int j = 1; j <= nj - 1; j++) {
            h2 = h2 + h1;
            wt = wt1 - wt;
            fj = wt;
            const double rj = rs[i0] + h2;
            for (int k = 1; k <= ip; k++) fj = fj * rj;
            gj = h2 * (dr - h2);
            double c1 = (cr[2] - cr[1]) * h2 + cr[0];
            double c2 = cr[3] * gj;
            gr = gr - fj * c1 * c2 / (1.0 + c2);
            c1 = (ci[2] - ci[1]) * h2 + ci[0];
            c2 = ci[3] * gj;
            gi = gi - fj * c1 * c2 / (1.0 + c2);
        }
        g[ip1 - 1] = g[i - 1] + 2.0 * dr * (gr + I * gi);
        if (i > 1 && ip2 < nr) {
            ip2 += 1;
            int t = i1;
            i1 = i2;
            i2 = i3;
            i3 = i4;
            i4 = t;
        }
        i0 = (i0 + 1) % 4;
    }
}

/* ---- one spin channel (zsl, zsr) of gf_local.f for mtasa = 1 / the sphere
 * part: int_zz_zj.f:116-165 (zlzl_l = int 2 z_l^L z_l^R, zljl_l = int
 * 2 z_l^L j_l^R when zj_flag), mdosms_c.f and mgreen_c.f, without the spin
 * factor.  tau: kkrsz x kkrsz column-major; zl_*, jl: iprpts x (lmax+1). */
static void mt_channel(int lmax, const zc *tau, const zc *zl_left, const zc *zl_right,
                       const zc *jl, int zj_flag, const double *r_mesh, int jmt, int jws,
                       double r_sph, int iprpts, zc *dos, zc *green, zc *zz_out, zc *zj_out)
{
    const int kkrsz = (lmax + 1) * (lmax + 1);
    const double pi = 3.14159265358979323846;
    double *rtmp = (double *)malloc(sizeof(double) * (size_t)(jmt + 1));
    double *work = (double *)malloc(sizeof(double) * (size_t)(jmt + 1));
    zc *fr = (zc *)malloc(sizeof(zc) * (size_t)(jmt + 1));
    zc *fint = (zc *)malloc(sizeof(zc) * (size_t)(jmt + 1));
    zc zzl[16];
    rtmp[0] = 0.0;
    for (int ir = 1; ir <= jmt; ir++) rtmp[ir] = sqrt(r_mesh[ir - 1]);
    const double xs = sqrt(r_sph);
    zc zzsum = 0.0, zjsum = 0.0;
    for (int l = 0; l <= lmax; l++) {
        const zc *zL = zl_left + (size_t)l * iprpts, *zR = zl_right + (size_t)l * iprpts;
        const zc *j = jl + (size_t)l * iprpts;
        for (int ir = 1; ir <= jmt; ir++) fr[ir] = 2.0 * zL[ir - 1] * zR[ir - 1];
        fr[0] = fr[1];
        or_cnewint(jmt + 1, rtmp, fr, fint, 5);
        const zc zlzl = cinterp_value(rtmp, fint, jmt + 1, xs, work);
        zc zljl = 0.0;
        if (zj_flag) {
            for (int ir = 1; ir <= jmt; ir++) fr[ir] = 2.0 * zL[ir - 1] * j[ir - 1];
            fr[0] = fr[1];
            or_cnewint(jmt + 1, rtmp, fr, fint, 5);
            zljl = cinterp_value(rtmp, fint, jmt + 1, xs, work);
        }
        zzl[l] = zlzl;
        if (zz_out) zz_out[l] = zlzl;
        if (zj_out) zj_out[l] = zljl;
        for (int m = -l; m <= l; m++) zjsum += zljl;
    }
    /* mdosms: sum over the kkrsz^2 entries of the diagonal pzzck times tau */
    int lm = 0;
    for (int l = 0; l <= lmax; l++)
        for (int m = -l; m <= l; m++, lm++) zzsum += zzl[l] * tau[lm + (size_t)lm * kkrsz];
    zc d = zzsum;
    if (zj_flag) d = d - zjsum;
    *dos = -d / pi;
    if (green) {
        for (int ir = 0; ir < jws; ir++) green[ir] = 0.0;
        int i = kkrsz - 1;
        for (int l = lmax; l >= 0; l--) {
            zc wt = 0.0;
            for (int m = -l; m <= l; m++) wt += tau[i + (size_t)i * kkrsz], i--;
            const zc *zL = zl_left + (size_t)l * iprpts, *zR = zl_right + (size_t)l * iprpts;
            const zc *j = jl + (size_t)l * iprpts;
            for (int ir = 0; ir < jws; ir++) {
                zc t = zR[ir] * wt;
                if (zj_flag) t = zR[ir] * wt - j[ir] * (double)(2 * l + 1);
                green[ir] += zL[ir] * t;
            }
        }
    }
    free(rtmp);
    free(work);
    free(fr);
    free(fint);
}

/* ---- the muffin-tin (ASA: whole-cell) part of green_function.f for one
 * collinear spin channel: gf_local.f:109-189 with int_zz_zj.f:116-165
 * (zlzl_l, zljl_l over the sphere, integrated in x = sqrt(r)), mdosms_c.f
 * (dos = -(sum zz tau - zj)/pi) and mgreen_c.f (green(r) for r < jws), with
 * the factor 2 of green_function.f:184-191 when n_spin_pola = 1.
 * tau00 kkrsz x kkrsz column-major; zlr, jlr iprpts x (lmax+1). */
int or_green_mt(int lmax, const zc *tau00, const zc *zlr, const zc *jlr, const double *r_mesh,
                int jmt, int jws, double r_sph, int iprpts, int n_spin_pola, zc *dos, zc *green,
                zc *zz_out, zc *zj_out)
{
    if (jmt < 4 || jmt > iprpts || jws > iprpts || lmax < 0 || lmax > 3) return -1;
    zc d;
    mt_channel(lmax, tau00, zlr, zlr, jlr, 1, r_mesh, jmt, jws, r_sph, iprpts, &d, green,
               zz_out, zj_out);
    const double spin = (n_spin_pola == 1) ? 2.0 : 1.0;
    *dos = spin * d;
    if (green)
        for (int ir = 0; ir < jws; ir++) green[ir] *= spin;
    return 0;
}

/* ---- spin-canted muffin-tin part of green_function.f:84-148 (isp = 1..4:
 * (zsl, zsr) = (1,1), (2,1), (1,2), (2,2), zj only on the diagonal) and the
 * recombination with wx, wy, wz (u_sigma_u) of :150-180.
 * tau00_l: 4 blocks kkrsz x kkrsz (trgtol's ml(nr,nc,4)); zlr, jlr:
 * iprpts x (lmax+1) x 2 spins; w = wx[4], wy[4], wz[4];
 * dos[4] = Tr, x, y, z; green[4*jws]: the same per r (may be NULL). */
int or_green_mt_canted(int lmax, const zc *tau00_l, const zc *zlr, const zc *jlr,
                       const double *r_mesh, int jmt, int jws, double r_sph, int iprpts,
                       const zc *w, zc *dos, zc *green)
{
    if (jmt < 4 || jmt > iprpts || jws > iprpts || lmax < 0 || lmax > 3) return -1;
    const int kkrsz = (lmax + 1) * (lmax + 1);
    const size_t sp = (size_t)iprpts * (lmax + 1);
    static const int zsl[4] = {0, 1, 0, 1}, zsr[4] = {0, 0, 1, 1}, flag[4] = {1, 0, 0, 1};
    zc d[4];
    zc *g = green ? (zc *)malloc(sizeof(zc) * 4 * (size_t)(jws > 0 ? jws : 1)) : NULL;
    for (int isp = 0; isp < 4; isp++)
        mt_channel(lmax, tau00_l + (size_t)isp * kkrsz * kkrsz, zlr + zsl[isp] * sp,
                   zlr + zsr[isp] * sp, jlr + zsr[isp] * sp, flag[isp], r_mesh, jmt, jws, r_sph,
                   iprpts, &d[isp], g ? g + (size_t)isp * jws : NULL, NULL, NULL);
    const zc *wx = w, *wy = w + 4, *wz = w + 8;
    dos[0] = d[0] + d[3];
    dos[1] = d[0] * wx[0] + d[1] * wx[2] + d[2] * wx[1] + d[3] * wx[3];
    dos[2] = d[0] * wy[0] + d[1] * wy[2] + d[2] * wy[1] + d[3] * wy[3];
    dos[3] = d[0] * wz[0] + d[1] * wz[2] + d[2] * wz[1] + d[3] * wz[3];
    if (g) {
        for (int ir = 0; ir < jws; ir++) {
            const zc g1 = g[ir], g2 = g[jws + ir], g3 = g[2 * (size_t)jws + ir],
                     g4 = g[3 * (size_t)jws + ir];
            green[ir] = g1 + g4;
            green[jws + ir] = g1 * wx[0] + g2 * wx[2] + g3 * wx[1] + g4 * wx[3];
            green[2 * (size_t)jws + ir] = g1 * wy[0] + g2 * wy[2] + g3 * wy[1] + g4 * wy[3];
            green[3 * (size_t)jws + ir] = g1 * wz[0] + g2 * wz[2] + g3 * wz[1] + g4 * wz[3];
        }
        free(g);
    }
    return 0;
}

/* ---- the interstitial part of the cell integrals (VORPOL/inter.f,
 * VORPOL/interstitial.f:1-162): for muffin-tin potentials (mtasa != 1)
 * int_zz_zj.f:244 adds, over the region between the muffin-tin sphere and
 * the circumscribing sphere of the Voronoi cell,
 *   pzz(L2,L1) = sum_points w Z_L1 Z_L2 sum_l3 cgnt(l3/2, L2, L1) sigma_l3,m1-m2(r)
 *   pzj        = sum_points w sum_L1 Z_L1 J_L1 sum_l3 cgnt(l3/2, L1, L1) sigma_l3,0(r)
 * with Z_l = matom_l j_l(pr) - (i e^{ipr}/r) h_l(pr), J_l = (i e^{ipr}/r)
 * h_l(pr) / matom_l (zsphbesjh: h_l scaled by e^{ix}/x), the cell's step
 * function sigma_lm(r) from wylm (m >= 0 stored, sigma_{l,-m} = (-1)^m
 * conj(sigma_lm)), the muffin-tin sphere taken out of sigma_00 (by sqrt(4 pi)
 * inside r_sph, half of it on r_sph to 1e-10), and Gauss points grwylm with
 * weights gwwylm per region (points with r <= 0 skipped).
 *   cgnt:  Gaunt table cgnt[l3/2 + (lmax_cg+1)(j1 + kkr_cg j2)], kkr_cg = (lmax_cg+1)^2
 *   grwylm, gwwylm: ngaussr x (ncrit-1); wylm: (2lmax+1)(lmax+1) x ngaussr x (ncrit-1)
 *   pzz: (lmax+1)^2 x (lmax+1)^2 column-major, pzz(L2, L1) at [L2 + kkr L1]. */
void or_interstitial(int zj_flag, int lmax, zc pnrel, const zc *matom_left,
                     const zc *matom_right, double r_sph, int ncrit, int ngaussr,
                     const double *grwylm, const double *gwwylm, const zc *wylm,
                     const double *cgnt, int lmax_cg, zc *pzz, zc *pzj)
{
    const int kkr = (lmax + 1) * (lmax + 1), kkr_cg = (lmax_cg + 1) * (lmax_cg + 1);
    const int nw = (2 * lmax + 1) * (lmax + 1);
    const double pi = 3.141592653589793, rt4pi = sqrt(4.0 * pi);
    zc bj[16], bh[16], eiz;
    for (int e = 0; e < kkr * kkr; e++) pzz[e] = 0.0;
    *pzj = 0.0;
    for (int ir = 0; ir < ncrit - 1; ir++)
        for (int ng = 0; ng < ngaussr; ng++) {
            const double rg = grwylm[ng + (long)ngaussr * ir];
            if (!(rg > 0.0)) continue;
            const double w = gwwylm[ng + (long)ngaussr * ir];
            const zc *wy = wylm + (long)nw * (ng + (long)ngaussr * ir);
            sph_bessel_jh(lmax, pnrel * rg, bj, bh, &eiz);
            eiz = I * eiz / rg;
            int lm1 = 0;
            for (int l1 = 0; l1 <= lmax; l1++) {
                const zc zlm1 = matom_left[l1] * bj[l1] - eiz * bh[l1];
                const zc jlm1 = eiz * bh[l1] / matom_left[l1];
                for (int m1 = -l1; m1 <= l1; m1++, lm1++) {
                    zc zljl = 0.0;
                    int lm2 = 0;
                    for (int l2 = 0; l2 <= lmax; l2++) {
                        const zc zlm2 = matom_right[l2] * bj[l2] - eiz * bh[l2];
                        for (int m2 = -l2; m2 <= l2; m2++, lm2++) {
                            zc zlzl = 0.0;
                            const int m3 = m1 - m2, am3 = m3 < 0 ? -m3 : m3;
                            const int dl = l1 > l2 ? l1 - l2 : l2 - l1;
                            const int llow = am3 > dl ? am3 : dl;
                            for (int l3 = l1 + l2; l3 >= llow; l3 -= 2) {
                                const zc wv = wy[(l3 * (l3 + 1)) / 2 + am3];
                                zc step = (m3 >= 0) ? wv : (double)(1 - 2 * (am3 % 2)) * conj(wv);
                                if (l3 == 0) {
                                    if (fabs(rg - r_sph) < 1e-10) step -= 0.5 * rt4pi;
                                    else if (rg < r_sph) step -= rt4pi;
                                }
                                const double c = cgnt[l3 / 2 + (lmax_cg + 1) * (lm2 + kkr_cg * lm1)];
                                zlzl += c * step;
                                if (lm2 == lm1 && zj_flag == 1) zljl += c * step;
                            }
                            pzz[lm2 + kkr * lm1] += zlzl * zlm1 * zlm2 * w;
                        }
                    }
                    *pzj += zljl * zlm1 * jlm1 * w;
                }
            }
        }
}
