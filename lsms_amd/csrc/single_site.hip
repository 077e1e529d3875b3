// single_site.hip -- single-site scattering on the GPU: the t-matrix
// t_l = 1/matom_l and the radial functions zlr, jlr of non-relativistic and
// scalar-relativistic spherical potentials, for many (energy, potential)
// pairs at once.
//
// Same algorithm as the reference's single_scatterer_nonrel
// (SingleSite/single_scatterer_nonrel.f:4-122) through semrel with solver 2
// (SingleSite/semrel_c.f:130-220): the potential r v(r) is fitted per mesh
// interval to a rational form (Misc/fitpot.f), the radial equation is
// integrated outwards from the first mesh point with an adaptive
// Bulirsch-Stoer step per interval (Misc/rwave.f, initwave.f, dfv_new.f,
// bulirschStoerIntegrator.cpp), matched at r_sph to j_l + alpha h_l
// (Misc/cinterp.f, fit.f, zsphbesjh.f, zsphbesj.f), and the irregular
// solution integrated inwards from r_sph.
//
// GPU mapping.  The work is a few hundred independent ODE solves of a few
// hundred mesh intervals each: latency-bound scalar FP64, far from any HBM or
// MFMA roofline.  k_fitpot gives one thread to each mesh interval;
// k_single_site gives one thread to each (energy, potential, l) and keeps the
// whole extrapolation tableau in registers (its loops are unrolled over the
// fixed step sequence, so no per-thread array is indexed at run time).  Each
// thread writes its own zlr column directly; the regular solution is kept
// raw in the zlr column until the matching gives its normalisation.  When
// jlr is asked for, a second set of threads runs the inward integrations at
// the same time, and k_jlr_scale applies their 1/matom factor afterwards.
#include <hip/hip_runtime.h>

#include <cmath>

#include "lsmsx_device.hpp"

namespace lsmsx {

namespace {

struct zd {
  double x, y;
};
__device__ inline zd mk(double a, double b = 0.0) { return zd{a, b}; }
__device__ inline zd operator+(zd a, zd b) { return zd{a.x + b.x, a.y + b.y}; }
__device__ inline zd operator-(zd a, zd b) { return zd{a.x - b.x, a.y - b.y}; }
__device__ inline zd operator-(zd a) { return zd{-a.x, -a.y}; }
__device__ inline zd operator*(zd a, zd b) { return zd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline zd operator*(double s, zd a) { return zd{s * a.x, s * a.y}; }
__device__ inline zd operator*(zd a, double s) { return zd{s * a.x, s * a.y}; }
__device__ inline zd operator/(zd a, double s) { return zd{a.x / s, a.y / s}; }
// complex quotient with Smith's scaling (no overflow of |b|^2)
__device__ inline zd operator/(zd a, zd b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return zd{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return zd{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}
__device__ inline zd operator/(double s, zd b) { return mk(s) / b; }
__device__ inline zd times_i(zd a) { return zd{-a.y, a.x}; }
// principal square root
__device__ inline zd zsqrt(zd z) {
  const double m = hypot(z.x, z.y);
  if (m == 0.0) return mk(0.0);
  if (z.x >= 0.0) {
    const double t = sqrt(0.5 * (m + z.x));
    return zd{t, z.y / (2.0 * t)};
  }
  const double t = sqrt(0.5 * (m - z.x));
  return zd{fabs(z.y) / (2.0 * t), copysign(t, z.y)};
}
__device__ inline zd zsin(zd z) { return zd{sin(z.x) * cosh(z.y), cos(z.x) * sinh(z.y)}; }
__device__ inline zd zcos(zd z) { return zd{cos(z.x) * cosh(z.y), -sin(z.x) * sinh(z.y)}; }
__device__ inline zd zexp_i(zd z) {  // e^{i z}
  const double s = exp(-z.y);
  return zd{s * cos(z.x), s * sin(z.x)};
}

// Bulirsch-Stoer step-count sequence n_j = 2 n_{j-2} (bulirschStoerIntegrator.cpp:66-67)
__constant__ int kSeq[12] = {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128};
constexpr int kBsSteps = 11;
constexpr int kBsTarget = 7;
constexpr double kBsEps = 1.0e-12;

// a/b from the hardware reciprocal refined by two Newton steps and one
// residual correction (about 1 ulp for the well-scaled operands of the
// radial equation); the IEEE division sequence is several dependent
// instructions longer, and these divisions sit on each thread's latency chain
__device__ inline double fdiv(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  const double q = a * r;
  return fma(r, fma(-b, q, a), q);
}

struct RhsPar {
  double er, ei, b, allp1;
};

// derivatives of (Re g, Im g, Re f, Im f) (dfv_new.f, nv = 4, eunit = 1)
__device__ inline void rhs4(double r, const double *y, double *dy, double rn, const double *pc,
                            const RhsPar &p) {
  const double s1or = fdiv(1.0, r), s1or2 = s1or * s1or;
  const double r1 = r - rn;
  const double v = fdiv(pc[0] * r1 + pc[1], 1.0 + pc[2] * r1 * r1) + pc[3];
  const double s2vmer = v * s1or - p.er;
  const double frelr = 1.0 - p.b * s2vmer;
  const double s2vmei = -p.ei;
  const double freli = -p.b * s2vmei;
  const double den = fdiv(p.allp1, frelr * frelr + freli * freli) * s1or2;
  const double facr = den * frelr + s2vmer;
  const double faci = -den * freli + s2vmei;
  dy[0] = frelr * y[2] - freli * y[3] + s1or * y[0];
  dy[1] = freli * y[2] + frelr * y[3] + s1or * y[1];
  dy[2] = facr * y[0] - faci * y[1] - y[2] * s1or;
  dy[3] = faci * y[0] + facr * y[1] - y[3] * s1or;
}

// modified midpoint rule with `steps` substeps over [x0, x1]
__device__ inline void midpoint4(double x0, double x1, const double *y0, double *y1, int steps,
                                 double rn, const double *pc, const RhsPar &p) {
  double y2[4], dy[4];
  const double h = fdiv(x1 - x0, (double)steps), h2 = 2.0 * h;
  double x = x0;
  rhs4(x, y0, dy, rn, pc, p);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    y1[i] = y0[i];
    y2[i] = y0[i] + h * dy[i];
  }
  for (int j = 1; j < steps; j++) {
    x += h;
    rhs4(x, y2, dy, rn, pc, p);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const double t = y1[i] + h2 * dy[i];
      y1[i] = y2[i];
      y2[i] = t;
    }
  }
  x += h;
  rhs4(x, y2, dy, rn, pc, p);
#pragma unroll
  for (int i = 0; i < 4; i++) y1[i] = 0.5 * (y1[i] + y2[i] + h * dy[i]);
}

// Adaptive Bulirsch-Stoer over [x0, x1] with polynomial extrapolation in h^2,
// one tableau row kept in registers (T[i][k] needs only T[i][k-1] and
// T[i-1][k-1]).  y is advanced in place; dy returns the derivative at x1.
// Returns 1 if the step size underflowed (y then holds the last estimate).
__device__ int bs_interval(double x0, double x1, double *y, double *dy, double rn,
                           const double *pc, const RhsPar &p) {
  const double xeps = 1.0e-12 * fabs(x0);
  double x = x0;
  const double h0 = x1 - x0;
  double h = h0;
  double row[kBsSteps][4] = {};
  double hh[kBsSteps] = {};
  int last = 0;  // diagonal index of the last tableau
  int rc = 0;
  do {
    double err = HUGE_VAL;
    int step = 0;
#pragma unroll
    for (int i = 0; i < kBsSteps; i++) {
      if (err > kBsEps) {
        double y1[4];
        midpoint4(x, x + h, y, y1, kSeq[i], rn, pc, p);
        hh[i] = fdiv(h, (double)kSeq[i]);
        double e = 0.0, mag = 0.0;
#pragma unroll
        for (int d = 0; d < 4; d++) {
          double old = row[0][d];  // T[i-1][0]
          row[0][d] = y1[d];
#pragma unroll
          for (int k = 1; k <= i; k++) {
            const double rr = fdiv(hh[i - k], hh[i]);
            const double den = fdiv(1.0, rr * rr - 1.0);
            const double keep = row[k][d];  // T[i-1][k] (unused when k == i)
            row[k][d] = row[k - 1][d] + (row[k - 1][d] - old) * den;
            if (k == i) e += fabs(row[i][d] - old);  // old = T[i-1][i-1]
            old = keep;
          }
          mag += fabs(row[i][d]);
        }
        if (i == 0) e = 4.0;
        if (mag == 0.0) mag = 1.0e-15;
        err = fdiv(e, mag);
        step = i + 1;
        last = i;
      }
    }
    if (err < kBsEps) {
      x = x + h;
      if (step < kBsTarget)
        h *= 1.25;
      else
        h *= (double)kSeq[kBsTarget] / (double)kSeq[step];
      const double lim = fabs(x1 - x);
      h = copysign(fabs(h) < lim ? fabs(h) : lim, h0);
#pragma unroll
      for (int k = 0; k < kBsSteps; k++)
        if (k == last) {
#pragma unroll
          for (int d = 0; d < 4; d++) y[d] = row[k][d];
        }
    } else {
      h = 0.25 * h;
      if (h + x == x) {
        rc = 1;
        break;
      }
    }
  } while (fabs(x - x1) > xeps);
#pragma unroll
  for (int k = 0; k < kBsSteps; k++)
    if (k == last) {
#pragma unroll
      for (int d = 0; d < 4; d++) y[d] = row[k][d];
    }
  rhs4(x1, y, dy, rn, pc, p);
  return rc;
}

// j_l(x) and h_l(x) = hm[l] e^{ix}/x for l = 0..LB (zsphbesj.f, zsphbesjh.f job 1)
template <int LB>
__device__ void bessel_jh(zd x, zd *bj, zd *hm, zd &eiz) {
  zd bs[LB + 1];
  const zd as = zsin(x), ac = zcos(x), x1 = 1.0 / x;
  bj[0] = as * x1;
  int k = ((int)(0.75 * (fabs(x.x) + fabs(x.y))) + 2) >> 2;
  bs[0] = as;
  bool done = false;
  if (k >= 1) {
    if (k > LB) k = LB;
    bs[1] = bj[0] - ac;
    bj[1] = bs[1] * x1;
#pragma unroll
    for (int l = 2; l <= LB; l++)
      if (l <= k) {
        bs[l] = (double)(2 * l - 1) * bj[l - 1] - bs[l - 2];
        bj[l] = bs[l] * x1;
      }
    done = (k == LB);
  }
  if (!done) {
    zd a = bs[0];
#pragma unroll
    for (int l = 1; l <= LB; l++)
      if (l == k) a = bs[l];
    const int nm = LB << 4;
    zd aj = mk((double)(2 * nm + 3));
    const zd x2 = x * x;
    for (int l = nm; l >= LB + 2; l--) aj = mk((double)(2 * l + 1)) - x2 / aj;
    bs[LB] = (double)(2 * LB + 3) * aj * x1 - x;
    bj[LB] = bs[LB] * x1;
    bs[LB - 1] = (double)(2 * LB + 1) * bj[LB] - aj;
#pragma unroll
    for (int l = LB - 1; l >= 1; l--)
      if (l >= k + 1) {
        bj[l] = bs[l] * x1;
        bs[l - 1] = (double)(2 * l + 1) * bj[l] - bs[l + 1];
      }
    zd bk = bs[0];
#pragma unroll
    for (int l = 1; l <= LB; l++)
      if (l == k) bk = bs[l];
    aj = a / bk;
#pragma unroll
    for (int l = 1; l <= LB; l++)
      if (l >= k + 1) bj[l] = aj * bj[l];
  }
  eiz = zexp_i(x);
  hm[0] = mk(0.0, -1.0);
  hm[1] = mk(-(1.0 - x1.y), -x1.x);
#pragma unroll
  for (int l = 2; l <= LB; l++) hm[l] = (double)(2 * l - 1) * hm[l - 1] * x1 - hm[l - 2];
}

// j_l, j_{l-1} (or j_1 at l = 0) and the same of hm, at argument x
struct BesselPick {
  zd bj_l, bj_o, hm_l, hm_o, hm_0, bj_0, eiz;
};

template <int LB>
__device__ BesselPick bessel_pick_t(zd x, int l) {
  zd bj[LB + 1], hm[LB + 1], eiz;
  bessel_jh<LB>(x, bj, hm, eiz);
  BesselPick b;
  b.eiz = eiz;
  b.bj_0 = bj[0];
  b.hm_0 = hm[0];
  b.bj_l = bj[0];
  b.hm_l = hm[0];
  b.bj_o = bj[1];
  b.hm_o = hm[1];
#pragma unroll
  for (int q = 1; q <= LB; q++)
    if (q == l) {
      b.bj_l = bj[q];
      b.hm_l = hm[q];
      b.bj_o = bj[q - 1];
      b.hm_o = hm[q - 1];
    }
  return b;
}

__device__ BesselPick bessel_pick(zd x, int l) {
  if (l <= 1) return bessel_pick_t<1>(x, l);
  if (l == 2) return bessel_pick_t<2>(x, l);
  return bessel_pick_t<3>(x, l);
}

// regular-solution value and derivative factors outside the sphere
// (rwave.f:153-162 and :228-236): j_l, d(r j_l)-like term, h_l r, its derivative
__device__ inline void outside(const BesselPick &b, zd ka, zd kar, int l, zd &bj, zd &dbj,
                               zd &bh1, zd &dbh1) {
  bj = b.bj_l;
  bh1 = b.hm_l * b.eiz / ka;
  if (l == 0) {
    dbj = -(ka * (b.bj_o - b.bj_0 / kar));
    dbh1 = -((b.hm_o - b.hm_0 / kar) * b.eiz);
  } else {
    const zd cd = (double)l / kar;
    dbj = ka * (b.bj_o - b.bj_l * cd);
    dbh1 = (b.hm_o - b.hm_l * cd) * b.eiz;
  }
}

// four-point rational fit of real f at 0-based point i (fit.f); the points it
// reads are passed as values: f at i1, i, ip1, i2
__device__ void fit4(const double *r, int nr, int i, int i1, int i2, double fi1, double fi,
                     double fip1, double fi2, double *c) {
  const int ip1 = i + 1;
  const double dr = r[ip1] - r[i];
  c[2] = fdiv(fip1 - fi, dr);
  c[0] = fi;
  c[1] = 0.0;
  c[3] = 0.0;
  if (i1 != i2) {
    c[1] = fdiv(fi2 - fip1, r[i2] - r[ip1]) - fdiv(fi1 - fi, r[i1] - r[i]);
    if (c[1] * (fi2 - fi1) * (r[i2] - r[i1]) > 0.0) c[1] = -c[1];
    double h1 = r[i1] - r[i];
    double d0 = fi1 - fi;
    const double a1 = d0 - c[2] * h1, a2 = d0 - c[1] * h1;
    h1 = h1 * (r[ip1] - r[i1]);
    const double e12 = -h1 * a2;
    double h2 = r[i2] - r[i];
    d0 = fi2 - fi;
    const double a3 = d0 - c[2] * h2, a4 = d0 - c[1] * h2;
    h2 = h2 * (r[ip1] - r[i2]);
    const double e22 = -h2 * a4;
    double gj = (a4 - a2) * h1 * h2;
    if (gj == 0.0) {
      c[3] = 0.0;
    } else {
      c[0] = fdiv(a1 * e22 - a3 * e12, gj);
      c[3] = fdiv(a1 * h2 - a3 * h1, gj);
    }
    gj = c[3] * dr * dr;
    if (gj > -4.0 && fabs(gj) > 1e-14 && gj < 1e14) {
      c[0] = fdiv(c[0], c[3]);
    } else {
      c[0] = fi;
      c[3] = 0.0;
    }
  } else {
    // three-point form (only for meshes of fewer than five points)
    const double fj = 0.5 * (fi + fip1);
    c[0] = fi - fj;
    c[1] = 0.0;
    const double rj = r[i] + 0.5 * dr;
    int j1 = i1;
    double fj1 = fi1;
    if (fabs(r[i1] - rj) > fabs(r[i2] - rj)) {
      j1 = i2;
      fj1 = fi2;
    }
    double h1 = r[j1] - r[i];
    const double a1 = fj1 - fi - c[2] * h1;
    h1 = h1 * (r[ip1] - r[j1]);
    if (fj1 != fj) {
      c[3] = fdiv(-a1, (fj1 - fj) * h1);
      const double gj = c[3] * dr * dr;
      if (gj <= -4.0 || fabs(gj) <= 1e-14 || gj >= 1e14) {
        c[0] = fi;
        c[3] = 0.0;
      }
    } else {
      c[0] = fi;
      c[3] = 0.0;
    }
  }
}

// cinterp.f value at rs from the four captured complex samples
__device__ zd interp4(const double *r, int nr, int i, int i1, int i2, double rs, const zd *f4) {
  double part[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    double c[4];
    const double g1 = q ? f4[0].y : f4[0].x, g = q ? f4[1].y : f4[1].x;
    const double gp = q ? f4[2].y : f4[2].x, g2 = q ? f4[3].y : f4[3].x;
    fit4(r, nr, i, i1, i2, g1, g, gp, g2, c);
    const double h1 = rs - r[i], h2 = r[i + 1] - rs;
    const double den = 1.0 + c[3] * h1 * h2;
    const double ps = ((c[2] - c[1]) * h1 + c[0]) / den;
    part[q] = ps + g - c[0] + c[1] * h1;
  }
  return mk(part[0], part[1]);
}

}  // namespace

// fitpot.f (nr = 4) for every potential: one thread per mesh interval.
// coef[(p*iprpts + j)*4 + q]; intervals j >= iend-2 get zeros.
__global__ void k_fitpot(const double *__restrict__ rmesh, const double *__restrict__ vr,
                         const int *__restrict__ iend_of, int iprpts, int npot,
                         double *__restrict__ coef) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)npot * iprpts) return;
  const int p = (int)(t / iprpts), j = (int)(t % iprpts);
  const double *r = rmesh + (long)p * iprpts, *rv = vr + (long)p * iprpts;
  double *c = coef + 4 * t;
  const int nfit = iend_of[p] - 1;
  if (j >= nfit - 1) {
    c[0] = c[1] = c[2] = c[3] = 0.0;
    return;
  }
  const double r0 = r[j], dr = r[j + 1] - r0;
  const double drvdr = (rv[j + 1] - rv[j]) / dr;
  int i1 = (j == 0) ? 2 : j - 1;
  if (j == nfit - 2) i1 -= 1;
  double h1 = r[i1] - r0, h2 = r[i1] - r[j + 1];
  const double e11 = h2, e12 = -h1 * (rv[i1] - rv[j + 1]);
  const double rhs1 = (rv[i1] - rv[j]) / h1 - drvdr;
  i1 += 1;
  if (i1 == j) i1 += 2;
  h1 = r[i1] - r0;
  h2 = r[i1] - r[j + 1];
  const double e21 = h2, e22 = -h1 * (rv[i1] - rv[j + 1]);
  const double rhs2 = (rv[i1] - rv[j]) / h1 - drvdr;
  const double det = e11 * e22 - e21 * e12;
  if (fabs(det) > 1e-16) {
    double b = -(e12 * rhs2 - e22 * rhs1) / det;
    const double cc = (e11 * rhs2 - e21 * rhs1) / det;
    const double tt = cc * dr * dr;
    if (tt > -1.0 && fabs(tt) > 1e-14) {
      const double a4 = rv[j + 1] + b / cc;
      c[0] = drvdr - b * dr;
      c[1] = rv[j] - a4;
      c[2] = cc;
      c[3] = a4;
      return;
    }
  }
  c[0] = drvdr;
  c[1] = rv[j];
  c[2] = 0.0;
  c[3] = 0.0;
}

// One thread per (energy e, potential p, l): rwave.f with the outputs of
// semrel_c.f (solver 2).  matom[(e*npot + p)*(lmax+1) + l];
// zlr/jlr (may be null) columns of iprpts at ((e*npot + p)*(lmax+1) + l)*iprpts.
__global__ void __launch_bounds__(64) k_single_site(
    const double *__restrict__ rmesh, const double *__restrict__ coef,
    const SingleSitePot *__restrict__ pots, const double2 *__restrict__ energy,
    const double2 *__restrict__ prel, int npot, int ne, int lmax, int iprpts, int srel,
    double2 *__restrict__ matom_out, double2 *__restrict__ zlr, double2 *__restrict__ jlr,
    int *__restrict__ status) {
  // threads [0, ntask): outward integration, matching, zlr and the jlr
  // beyond the sphere; with jlr, threads [ntask, 2 ntask) run the inward
  // (irregular) integration of the same task at the same time, since it
  // starts from the Hankel function at r_sph and does not need the regular
  // solution.  k_jlr_scale divides their raw values by matom afterwards.
  const int nl = lmax + 1;
  const long ntask = (long)ne * npot * nl;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (jlr ? 2 : 1) * ntask) return;
  const bool inward = t >= ntask;
  if (inward) t -= ntask;
  const int l = (int)(t % nl), p = (int)((t / nl) % npot), e = (int)(t / ((long)nl * npot));
  const SingleSitePot P = pots[p];
  const double *rad = rmesh + (long)p * iprpts;
  const double *cf = coef + (long)p * iprpts * 4;
  const zd E = mk(energy[e].x, energy[e].y);
  const zd ka = zsqrt(E + mk(0.0, 1.0e-14));
  const double al = (double)l;
  const int jmt = P.jmt, jws = P.jws;
  double2 *zcol = zlr ? zlr + t * iprpts : nullptr;
  double2 *jcol = jlr ? jlr + t * iprpts : nullptr;
  int rc = 0;

  // initial values (initwave.f, core = false; r_1 > 0 so istart = 1)
  RhsPar rp;
  rp.er = E.x;
  rp.ei = E.y;
  double y[4], dy[4];
  {
    const double c = 137.0359895, c2 = c * c;
    double z0 = cf[1] + cf[3];
    z0 = (double)(int)(fabs(z0) + 0.1);
    const double z = 0.5 * z0;
    const bool rel = srel && !(z < 0.5);
    rp.allp1 = fabs((double)(l * (l + 1)));
    const double x = rad[0];
    zd g1, g2;
    if (rel) {
      rp.b = 0.25 / c2;
      g1 = mk(1.0);
      g2 = mk(2.0 * c2 * (sqrt(rp.allp1 + 1.0 - z * z / c2) - 1.0) / z);
    } else {
      rp.b = 0.0;
      zd v0 = 0.5 * (E - mk(cf[0]));
      const double v1 = 2.0 * 0.5 * cf[1] * cf[2];
      v0 = v0 - mk(0.5 * v1 * (rad[0] + rad[1]));
      const double alp1 = al + 1.0;
      const double a1 = -z / alp1;
      const zd a2 = (mk(-z * a1) - v0) / (2.0 * al + 3.0);
      const zd a3 = -((-a1) * (mk(z * z) - (3.0 * al + 4.0) * v0) + mk((2.0 * al + 3.0) * v1)) /
                    (3.0 * (al + 2.0) * (2.0 * al + 3.0));
      const double rl = pow(x, al);
      g1 = (x * rl) * (mk(1.0) + x * (mk(a1) + x * (a2 + x * a3)));
      g2 = rl * (mk(al) + x * (mk(alp1 * a1) + x * ((al + 2.0) * a2 + x * (al + 3.0) * a3)));
    }
    y[0] = g1.x;
    y[1] = g1.y;
    y[2] = g2.x;
    y[3] = g2.y;
  }
  if (inward) {  // irregular solution inwards from r_sph (rwave.f:181-223), raw
    const zd kar = ka * P.r_sph;
    const BesselPick b = bessel_pick(kar, l);
    const zd hl = b.hm_l * b.eiz / ka;
    zd dh;
    if (l == 0) {
      dh = -(b.hm_o * b.eiz);
    } else {
      const zd cd = (double)(l + 1) / kar;
      dh = (b.hm_o - b.hm_l * cd) * b.eiz;
    }
    y[0] = hl.x;
    y[1] = hl.y;
    y[2] = dh.x;
    y[3] = dh.y;
    rhs4(P.r_sph, y, dy, P.r_sph, cf + 4 * (jmt - 1), rp);
    for (int j = jmt - 2; j >= 0; j--) {
      const double rfrom = (j + 1 == jmt - 1) ? P.r_sph : rad[j + 1];
      rc |= bs_interval(rfrom, rad[j], y, dy, rad[j], cf + 4 * j, rp);
      if (j < jws) jcol[j] = make_double2(y[0], y[1]);
    }
    if (rc) atomicOr(status, 1);
    return;
  }
  rhs4(rad[0], y, dy, rad[0], cf, rp);

  // outward integration; the raw regular solution goes to the zlr column and
  // the four-point matching window is captured in registers
  zd wr[4], wd[4];  // rr, drr at i1, i, ip1, i2
  const int wi = P.ip1 - 1;
  auto capture = [&](int j) {
    const zd g = mk(y[0], y[1]), d = mk(dy[0], dy[1]);
    if (j == P.i1) { wr[0] = g; wd[0] = d; }
    if (j == wi) { wr[1] = g; wd[1] = d; }
    if (j == P.ip1) { wr[2] = g; wd[2] = d; }
    if (j == P.i2) { wr[3] = g; wd[3] = d; }
    if (zcol && j < jws) zcol[j] = make_double2(y[0], y[1]);
  };
  capture(0);
  for (int j = 1; j < jmt; j++) {
    rc |= bs_interval(rad[j - 1], rad[j], y, dy, rad[j - 1], cf + 4 * (j - 1), rp);
    capture(j);
  }

  // match to j_l + alpha h_l at r_sph (rwave.f:123-171)
  const double r_sph = P.r_sph;
  const zd drrsph = interp4(rad, jmt, wi, P.i1, P.i2, r_sph, wd);
  const zd rrsph = interp4(rad, jmt, wi, P.i1, P.i2, r_sph, wr);
  const zd dfi = drrsph / rrsph, fi = rrsph / r_sph;
  zd kar = ka * r_sph;
  zd bj, dbj, bh1, dbh1;
  BesselPick b = bessel_pick(kar, l);
  outside(b, ka, kar, l, bj, dbj, bh1, dbh1);
  const zd cdum = -((bj * dfi - dbj) / (bh1 * dfi - dbh1));
  const zd cnorm = (bj + cdum * bh1) / fi;
  const zd tatoml = r_sph * cdum;
  const zd matom = -(times_i(ka) / tatoml);
  matom_out[t] = make_double2(matom.x, matom.y);
  const zd iprel = times_i(mk(prel[e].x, prel[e].y));

  // zlr = (cnorm g) matom / r inside the sphere
  if (zcol) {
    const int jend = jmt - 1 < jws ? jmt - 1 : jws;
    for (int j = 0; j < jend; j++) {
      const double2 g = zcol[j];
      const zd gn = cnorm * mk(g.x, g.y);
      const zd z = (gn * matom) / rad[j];
      zcol[j] = make_double2(z.x, z.y);
    }
  }

  // outside the sphere: Bessel and Hankel forms (rwave.f:225-240)
  for (int j = jmt - 1; j < jws; j++) {
    kar = ka * rad[j];
    b = bessel_pick(kar, l);
    outside(b, ka, kar, l, bj, dbj, bh1, dbh1);
    if (zcol) {
      const zd g = bj * rad[j] + bh1 * tatoml;
      const zd z = (g * matom) / rad[j];
      zcol[j] = make_double2(z.x, z.y);
    }
    if (jcol) {
      const zd q = (iprel * bh1) / (rad[j] * matom);
      jcol[j] = make_double2(q.x, q.y);
    }
  }
  if (rc) atomicOr(status, 1);
}

// jlr = i prel g_irr / (r matom) (semrel_c.f:214-217) for the points the
// inward threads of k_single_site left raw (j < min(jmt-1, jws)); one
// thread per (task, mesh point)
__global__ void k_jlr_scale(const double *__restrict__ rmesh,
                            const SingleSitePot *__restrict__ pots,
                            const double2 *__restrict__ prel, const double2 *__restrict__ matom,
                            int npot, int ne, int lmax, int iprpts, double2 *__restrict__ jlr) {
  const int nl = lmax + 1;
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ntask = (long)ne * npot * nl;
  if (g >= ntask * iprpts) return;
  const long t = g / iprpts;
  const int j = (int)(g % iprpts);
  const int p = (int)((t / nl) % npot), e = (int)(t / ((long)nl * npot));
  const SingleSitePot P = pots[p];
  const int jend = P.jmt - 1 < P.jws ? P.jmt - 1 : P.jws;
  if (j >= jend) return;
  const zd iprel = times_i(mk(prel[e].x, prel[e].y));
  const zd m = mk(matom[t].x, matom[t].y);
  const double2 v = jlr[t * iprpts + j];
  const zd q = (iprel * mk(v.x, v.y)) / (rmesh[(long)p * iprpts + j] * m);
  jlr[t * iprpts + j] = make_double2(q.x, q.y);
}

// ---------------------------------------------------------------------------
// Muffin-tin (ASA: whole-cell) DOS and Green function of one collinear spin
// channel from tau00 and the single-site radial functions: gf_local.f with
// the sphere integrals of int_zz_zj.f:116-165 (cnewint.f on x = sqrt(r),
// cinterp.f at sqrt(r_sph)), mdosms_c.f and mgreen_c.f, times 2 when
// n_spin_pola = 1 (green_function.f:184-191).
//
// One 64-thread workgroup per eight solves (kGreenSolves): lane 8s+2l and
// 8s+2l+1 run the running integrals of 2 z_l z_l and 2 z_l j_l of solve s
// over the jmt+1 points of the x mesh (a ring of four fitted points, as
// cnewint keeps them), so every lane of the wave carries one sequential
// integral; lane s then contracts solve s's integrals with the diagonal of
// tau00, and all lanes form green(r) solve by solve, reading zlr/jlr with
// consecutive lanes on consecutive radial points.

namespace {

// running integral step of cnewint.f for one interval: slots of the ring in
// rs/fr/fi, interval from slot i0 to the next; returns the increment
__device__ zd cnewint_step(const double *rs, const double *fr, const double *fi, int i0, int ip) {
  const int i1 = (i0 + 3) & 3, ip1 = (i0 + 1) & 3, i2 = (i0 + 2) & 3;
  double xl[4], cr[4], ci[4];
  xl[0] = rs[i1];
  xl[1] = rs[i0];
  xl[2] = rs[ip1];
  xl[3] = rs[i2];
  fit4(xl, 4, 1, 0, 3, fr[i1], fr[i0], fr[ip1], fr[i2], cr);
  fit4(xl, 4, 1, 0, 3, fi[i1], fi[i0], fi[ip1], fi[i2], ci);
  const double x0 = rs[i0], rip1 = rs[ip1];
  const double pip1 = 1.0 / (double)(ip + 1), pip2 = 1.0 / (double)(ip + 2);
  double gj = 1.0, fj = rip1;
  for (int j = 1; j <= ip; j++) {
    gj = gj * x0 + fj;
    fj = fj * rip1;
  }
  fj = (gj * x0 + fj) * pip2;
  gj = gj * pip1;
  double gr = (fr[i0] - cr[2] * x0) * gj + cr[2] * fj;
  double gi = (fi[i0] - ci[2] * x0) * gj + ci[2] * fj;
  const double dr = rip1 - x0;
  const double h1 = dr / 6.0;
  const double wt0 = 2.0 / 18.0, wt1 = 3.0 * wt0;
  double h2 = 0.0, wt = wt0;
#pragma unroll
  for (int j = 1; j <= 5; j++) {
    h2 = h2 + h1;
    wt = wt1 - wt;
    fj = wt;
    const double rj = x0 + h2;
    for (int k = 1; k <= ip; k++) fj = fj * rj;
    gj = h2 * (dr - h2);
    double c1 = (cr[2] - cr[1]) * h2 + cr[0];
    double c2 = cr[3] * gj;
    gr = gr - fdiv(fj * c1 * c2, 1.0 + c2);
    c1 = (ci[2] - ci[1]) * h2 + ci[0];
    c2 = ci[3] * gj;
    gi = gi - fdiv(fj * c1 * c2, 1.0 + c2);
  }
  return mk(2.0 * dr * gr, 2.0 * dr * gi);
}

}  // namespace

namespace {

// int_zz_zj.f:116-165 for one lane: the running integral (cnewint.f, ip0 = 5,
// on x_0 = 0, x_k = sqrt(r_k)) of 2 a_l b_l with f_0 = f_1, interpolated at
// sqrt(r_sph) (cinterp.f); a, b are radial columns, rm the mesh
__device__ zd sphere_integral(const double2 *__restrict__ a, const double2 *__restrict__ bcol,
                              const double *__restrict__ rm, const GreenPot &P) {
  auto fval = [&](int k) -> zd {
    const int q = k == 0 ? 0 : k - 1;
    const zd u = mk(a[q].x, a[q].y), c = mk(bcol[q].x, bcol[q].y);
    return 2.0 * (u * c);
  };
  auto xval = [&](int k) -> double { return k == 0 ? 0.0 : sqrt(rm[k - 1]); };
  const int nr = P.jmt + 1;
  const int ip = 2 * 5 + 1;
  double rs[4], fr[4], fi[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    rs[k] = sqrt(xval(k));
    const zd f = fval(k);
    fr[k] = f.x;
    fi[k] = f.y;
  }
  zd g = mk(0.0), w[4];
  int i4 = 3, i0 = 0, ip2 = 4;
  auto capture = [&](int k) {
    if (k == P.i1) w[0] = g;
    if (k == P.ip1 - 1) w[1] = g;
    if (k == P.ip1) w[2] = g;
    if (k == P.i2) w[3] = g;
  };
  capture(0);
  for (int i = 1; i <= nr - 1; i++) {
    if (i > 2 && ip2 <= nr) {
      const zd f = fval(ip2 - 1);
      const double xv = sqrt(xval(ip2 - 1));
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (q == i4) {
          rs[q] = xv;
          fr[q] = f.x;
          fi[q] = f.y;
        }
    }
    g = g + cnewint_step(rs, fr, fi, i0, ip);
    capture(i);
    if (i > 1 && ip2 < nr) {
      ip2 += 1;
      i4 = (i4 + 1) & 3;  // the ring rotation i1..i4 advances i4 by one slot
    }
    i0 = (i0 + 1) & 3;
  }
  double xw[4];
  xw[0] = xval(P.i1);
  xw[1] = xval(P.ip1 - 1);
  xw[2] = xval(P.ip1);
  xw[3] = xval(P.i2);
  const int same = (P.i1 == P.i2);
  return interp4(xw, 4, 1, 0, same ? 0 : 3, P.xs, w);
}

}  // namespace

__global__ void __launch_bounds__(64) k_green_mt(
    const double2 *__restrict__ tau00, const double2 *__restrict__ zlr,
    const double2 *__restrict__ jlr, const double *__restrict__ rmesh,
    const GreenPot *__restrict__ pots, const int *__restrict__ pot_of,
    const int *__restrict__ src_of, int nsolve, int lmax, int iprpts, double spin,
    double2 *__restrict__ dos, double2 *__restrict__ green) {
  // zlr/jlr row of solve b: src_of[b] (radial functions of a single-site
  // batch left on the device) or b
  __shared__ zd s_int[kGreenSolves][2][4];
  const int tid = threadIdx.x;
  const int nl = lmax + 1, kk = nl * nl;
  const int b0 = blockIdx.x * kGreenSolves;
  // phase 1: lane 8s + 2l + which runs integral `which` of channel l of solve b0 + s
  {
    const int sl = tid >> 3, l = (tid & 7) >> 1, which = tid & 1;
    const int b = b0 + sl;
    if (b < nsolve && l < nl) {
      const GreenPot P = pots[pot_of[b]];
      const double *rm = rmesh + (long)pot_of[b] * iprpts;
      const long src = src_of ? src_of[b] : b;
      const double2 *zl = zlr + (src * nl + l) * iprpts;
      const double2 *ol = (which ? jlr : zlr) + (src * nl + l) * iprpts;
      s_int[sl][which][l] = sphere_integral(zl, ol, rm, P);
    }
  }
  __syncthreads();
  // phase 2: mdosms, -(sum pzzck tau - pzjck)/pi, one lane per solve
  if (tid < kGreenSolves && b0 + tid < nsolve) {
    const int b = b0 + tid;
    const double2 *tb = tau00 + (long)b * kk * kk;
    zd zz = mk(0.0), zj = mk(0.0);
    for (int l = 0; l < nl; l++)
      for (int m = -l; m <= l; m++) zj = zj + s_int[tid][1][l];
    int lm = 0;
    for (int l = 0; l < nl; l++)
      for (int m = -l; m <= l; m++, lm++) {
        const double2 t = tb[lm + (long)lm * kk];
        zz = zz + s_int[tid][0][l] * mk(t.x, t.y);
      }
    const zd d = -(zz - zj) / 3.14159265358979323846;
    dos[b] = make_double2(spin * d.x, spin * d.y);
  }
  // phase 3: mgreen, sum over l from lmax down of z (z wt_l - j (2l+1)),
  // each solve's radial points spread over the 64 lanes
  if (green) {
    for (int sl = 0; sl < kGreenSolves; sl++) {
      const int b = b0 + sl;
      if (b >= nsolve) break;
      const double2 *tb = tau00 + (long)b * kk * kk;
      zd wt[4];
      int i = kk - 1;
      for (int l = lmax; l >= 0; l--) {
        zd sum = mk(0.0);
        for (int m = -l; m <= l; m++, i--) {
          const double2 t = tb[i + (long)i * kk];
          sum = sum + mk(t.x, t.y);
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (q == l) wt[q] = sum;
      }
      const int jws = pots[pot_of[b]].jws;
      const long src = src_of ? src_of[b] : b;
      const double2 *z0 = zlr + src * nl * iprpts, *j0 = jlr + src * nl * iprpts;
      double2 *gb = green + (long)b * iprpts;
      for (int ir = tid; ir < jws; ir += blockDim.x) {
        zd acc = mk(0.0);
#pragma unroll
        for (int l = 3; l >= 0; l--) {
          if (l > lmax) continue;
          const double2 zv = z0[(long)l * iprpts + ir], jv = j0[(long)l * iprpts + ir];
          const zd zz = mk(zv.x, zv.y), jj = mk(jv.x, jv.y);
          acc = acc + zz * (zz * wt[l] - jj * (double)(2 * l + 1));
        }
        gb[ir] = make_double2(spin * acc.x, spin * acc.y);
      }
    }
  }
}

// Spin-canted muffin-tin DOS and Green function (green_function.f:84-180):
// the four spin channels isp = (zsl, zsr) = (1,1), (2,1), (1,2), (2,2), the
// irregular term only on the diagonal ones, then Tr and the x, y, z parts
// with wx, wy, wz (u_sigma_u.f).  Per solve: tau00_l as four kkrsz x kkrsz
// blocks (trgtol.f), zlr/jlr as 2 spins x (lmax+1) x iprpts, w = wx, wy, wz.
// One 64-thread workgroup per two solves: lane 32s + 8isp + 2l + which.
__global__ void __launch_bounds__(64) k_green_mt_canted(
    const double2 *__restrict__ tau00, const double2 *__restrict__ zlr,
    const double2 *__restrict__ jlr, const double2 *__restrict__ wxyz,
    const double *__restrict__ rmesh, const GreenPot *__restrict__ pots,
    const int *__restrict__ pot_of, int nsolve, int lmax, int iprpts,
    double2 *__restrict__ dos, double2 *__restrict__ green) {
  __shared__ zd s_int[2][4][2][4];
  __shared__ zd s_d[2][4];
  const int tid = threadIdx.x;
  const int nl = lmax + 1, kk = nl * nl;
  const int b0 = blockIdx.x * 2;
  {
    const int sl = tid >> 5, isp = (tid >> 3) & 3, l = (tid & 7) >> 1, which = tid & 1;
    const int b = b0 + sl;
    const bool diag = (isp == 0 || isp == 3);
    if (b < nsolve && l < nl && (which == 0 || diag)) {
      const int zsl = (isp & 1), zsr = (isp >> 1);
      const GreenPot P = pots[pot_of[b]];
      const double *rm = rmesh + (long)pot_of[b] * iprpts;
      const double2 *zl = zlr + (((long)b * 2 + zsl) * nl + l) * iprpts;
      const double2 *ol = (which ? jlr : zlr) + (((long)b * 2 + zsr) * nl + l) * iprpts;
      s_int[sl][isp][which][l] = sphere_integral(zl, ol, rm, P);
    }
  }
  __syncthreads();
  if (tid < 8) {  // mdosms per channel
    const int sl = tid >> 2, isp = tid & 3, b = b0 + sl;
    if (b < nsolve) {
      const bool diag = (isp == 0 || isp == 3);
      const double2 *tb = tau00 + ((long)b * 4 + isp) * kk * kk;
      zd zz = mk(0.0), zj = mk(0.0);
      for (int l = 0; l < nl; l++)
        for (int m = -l; m <= l; m++) zj = zj + (diag ? s_int[sl][isp][1][l] : mk(0.0));
      int lm = 0;
      for (int l = 0; l < nl; l++)
        for (int m = -l; m <= l; m++, lm++) {
          const double2 t = tb[lm + (long)lm * kk];
          zz = zz + s_int[sl][isp][0][l] * mk(t.x, t.y);
        }
      zd d = zz;
      if (diag) d = d - zj;
      s_d[sl][isp] = -d / 3.14159265358979323846;
    }
  }
  __syncthreads();
  if (tid < 2 && b0 + tid < nsolve) {  // recombination (green_function.f:150-162)
    const int b = b0 + tid;
    const zd d1 = s_d[tid][0], d2 = s_d[tid][1], d3 = s_d[tid][2], d4 = s_d[tid][3];
    zd out[4];
    out[0] = d1 + d4;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double2 *w = wxyz + (long)b * 12 + 4 * a;
      out[1 + a] = d1 * mk(w[0].x, w[0].y) + d2 * mk(w[2].x, w[2].y) +
                   d3 * mk(w[1].x, w[1].y) + d4 * mk(w[3].x, w[3].y);
    }
#pragma unroll
    for (int a = 0; a < 4; a++) dos[(long)b * 4 + a] = make_double2(out[a].x, out[a].y);
  }
  if (green) {  // mgreen per channel, then the same recombination per r
    for (int sl = 0; sl < 2; sl++) {
      const int b = b0 + sl;
      if (b >= nsolve) break;
      zd wt[4][4];
#pragma unroll
      for (int isp = 0; isp < 4; isp++) {
        const double2 *tb = tau00 + ((long)b * 4 + isp) * kk * kk;
        int i = kk - 1;
        for (int l = lmax; l >= 0; l--) {
          zd sum = mk(0.0);
          for (int m = -l; m <= l; m++, i--) {
            const double2 t = tb[i + (long)i * kk];
            sum = sum + mk(t.x, t.y);
          }
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (q == l) wt[isp][q] = sum;
        }
      }
      zd w[3][4];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const double2 v = wxyz[(long)b * 12 + 4 * a + q];
          w[a][q] = mk(v.x, v.y);
        }
      const int jws = pots[pot_of[b]].jws;
      const double2 *zb = zlr + (long)b * 2 * nl * iprpts, *jb = jlr + (long)b * 2 * nl * iprpts;
      for (int ir = tid; ir < jws; ir += blockDim.x) {
        zd g[4];
#pragma unroll
        for (int isp = 0; isp < 4; isp++) {
          const int zsl = (isp & 1), zsr = (isp >> 1);
          const bool diag = (isp == 0 || isp == 3);
          zd acc = mk(0.0);
#pragma unroll
          for (int l = 3; l >= 0; l--) {
            if (l > lmax) continue;
            const double2 zL = zb[((long)zsl * nl + l) * iprpts + ir];
            const double2 zR = zb[((long)zsr * nl + l) * iprpts + ir];
            const double2 jR = jb[((long)zsr * nl + l) * iprpts + ir];
            zd t = mk(zR.x, zR.y) * wt[isp][l];
            if (diag) t = mk(zR.x, zR.y) * wt[isp][l] - mk(jR.x, jR.y) * (double)(2 * l + 1);
            acc = acc + mk(zL.x, zL.y) * t;
          }
          g[isp] = acc;
        }
        double2 *gb = green + (long)b * 4 * iprpts;
        const zd g0 = g[0] + g[3];
        gb[ir] = make_double2(g0.x, g0.y);
#pragma unroll
        for (int a = 0; a < 3; a++) {
          const zd v = g[0] * w[a][0] + g[1] * w[a][2] + g[2] * w[a][1] + g[3] * w[a][3];
          gb[(long)(1 + a) * iprpts + ir] = make_double2(v.x, v.y);
        }
      }
    }
  }
}

// t-matrix store from a single-site batch: block (slot e, spin isp) of store
// column s = kkr x kkr diagonal 1/matom_l (single_scatterer_nonrel.f:100-111)
// at [s*ldT + e*blk + isp*kkr^2], the collinear layout the path reads.
// One thread per block entry; the caller zeroes the store first.
__global__ void k_tmat_fill(const double2 *__restrict__ matom, const int *__restrict__ store_of,
                            const int *__restrict__ spin_of, int ne, int npot, int lmax,
                            long ldT, long blk, double2 *__restrict__ store) {
  const int nl = lmax + 1, kkr = nl * nl;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ne * npot * kkr * kkr) return;
  const int idx = (int)(t % (kkr * kkr));
  const long q = t / (kkr * kkr);
  const int p = (int)(q % npot), e = (int)(q / npot);
  const int i = idx % kkr, j = idx / kkr;
  double2 v = make_double2(0.0, 0.0);
  if (i == j) {
    int l = 0;
    while ((l + 1) * (l + 1) <= i) l++;
    const double2 m = matom[q * nl + l];
    const zd tv = 1.0 / mk(m.x, m.y);
    v = make_double2(tv.x, tv.y);
  }
  store[(long)store_of[p] * ldT + (long)e * blk + (long)spin_of[p] * kkr * kkr + idx] = v;
}

// Spin-canted t-matrix store from a single-site batch: the global-frame
// t_g = trltog(t_up, t_dn) of each atom (Misc/trltog.f, as
// calculateSingleScattererSolution, SingleSiteScattering.cpp:44-46): block
// (ri, ci) of the 2kkr x 2kkr store block is ud[ri] u[2ci] t_up +
// ud[2+ri] u[2ci+1] t_dn, diagonal; stored at [s*ldT + e*blk + i + j*2kkr].
// One thread per entry; the caller zeroes the store first.
__global__ void k_tmat_fill_canted(const double2 *__restrict__ matom,
                                   const int *__restrict__ atom_map, const double2 *__restrict__ uu,
                                   int ne, int npot, int natom, int lmax, long ldT, long blk,
                                   double2 *__restrict__ store) {
  const int nl = lmax + 1, kkr = nl * nl, n2 = 2 * kkr;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ne * natom * n2 * n2) return;
  const int idx = (int)(t % (n2 * n2));
  const long q = t / (n2 * n2);
  const int a = (int)(q % natom), e = (int)(q / natom);
  const int i = idx % n2, j = idx / n2;
  const int ri = i / kkr, ci = j / kkr, ii = i % kkr, jj = j % kkr;
  double2 v = make_double2(0.0, 0.0);
  if (ii == jj) {
    int l = 0;
    while ((l + 1) * (l + 1) <= ii) l++;
    const int s_col = atom_map[3 * a], p_up = atom_map[3 * a + 1], p_dn = atom_map[3 * a + 2];
    const double2 mu = matom[((long)e * npot + p_up) * nl + l];
    const double2 md = matom[((long)e * npot + p_dn) * nl + l];
    const zd tu = 1.0 / mk(mu.x, mu.y), td = 1.0 / mk(md.x, md.y);
    const double2 *u = uu + (long)a * 8, *ud = u + 4;
    const zd ca = mk(ud[ri].x, ud[ri].y) * mk(u[2 * ci].x, u[2 * ci].y);
    const zd cb = mk(ud[2 + ri].x, ud[2 + ri].y) * mk(u[2 * ci + 1].x, u[2 * ci + 1].y);
    const zd g = ca * tu + cb * td;
    v = make_double2(g.x, g.y);
    store[(long)s_col * ldT + (long)e * blk + i + (long)j * n2] = v;
    return;
  }
  store[(long)atom_map[3 * a] * ldT + (long)e * blk + i + (long)j * n2] = v;
}

// Interstitial part of the cell integrals of muffin-tin potentials
// (VORPOL/inter.f -> interstitial.f, added by int_zz_zj.f:244 for mtasa != 1)
// and its density of states, one workgroup per (atom, energy) solve b:
// thread e = lm1 kkr + lm2 accumulates pzz(lm2, lm1) over the cell's radial
// Gauss points (region-major, as interstitial.f), the diagonal threads also
// their pzj part; the spherical Bessel functions of each point are formed
// once (zsphbesjh, as the single-site matching) and shared through LDS.  Then
//   dos_b = -scale (sum_e pzz_e tau00_e - pzj) / pi      (mdosms_c.f)
// with fixed-order sums.  Cell c = cell_of[b]: ncrit[c] critical radii,
// grwylm/gwwylm [c][ir][ng] (ngaussr x maxreg per cell), wylm [c][ir][ng][nw].
template <int LB>
__device__ void interstitial_solve(int b, int zj_flag, const double2 *__restrict__ prel,
                                   const double2 *__restrict__ mleft,
                                   const double2 *__restrict__ mright,
                                   const double *__restrict__ r_sph, const int *__restrict__ cell_of,
                                   const int *__restrict__ ncrit, int ngaussr, int maxreg,
                                   const double *__restrict__ grwylm,
                                   const double *__restrict__ gwwylm,
                                   const double2 *__restrict__ wylm,
                                   const double *__restrict__ cgnt, int lmax_cg,
                                   const double2 *__restrict__ tau00, double scale,
                                   double2 *__restrict__ pzz_out, double2 *__restrict__ pzj_out,
                                   double2 *__restrict__ dos_out) {
  constexpr int kk = (LB + 1) * (LB + 1), nw = (2 * LB + 1) * (LB + 1);
  __shared__ zd sbj[LB + 1], sbh[LB + 1], seiz;
  __shared__ double red[2][256];
  __shared__ zd szj[kk];
  const int e = threadIdx.x, lm1 = e / kk, lm2 = e % kk;
  const bool act = e < kk * kk;
  const int l1 = act ? (int)sqrt((double)lm1 + 0.5) : 0, l2 = act ? (int)sqrt((double)lm2 + 0.5) : 0;
  const int m1 = lm1 - l1 * (l1 + 1), m2 = lm2 - l2 * (l2 + 1);
  const int c = cell_of[b], nreg = ncrit[c] - 1;
  const double rs = r_sph[c], rt4pi = sqrt(4.0 * 3.141592653589793);
  const int kkr_cg = (lmax_cg + 1) * (lmax_cg + 1);
  const zd p = mk(prel[b].x, prel[b].y);
  const zd ml = act ? mk(mleft[(long)b * (LB + 1) + l1].x, mleft[(long)b * (LB + 1) + l1].y) : mk(1.0);
  const zd mr = act ? mk(mright[(long)b * (LB + 1) + l2].x, mright[(long)b * (LB + 1) + l2].y) : mk(1.0);
  zd acc = mk(0.0), zj = mk(0.0);
  for (int ir = 0; ir < nreg; ir++)
    for (int ng = 0; ng < ngaussr; ng++) {
      const long pt = ((long)c * maxreg + ir) * ngaussr + ng;
      const double rg = grwylm[pt];
      if (!(rg > 0.0)) continue;  // uniform across the workgroup
      __syncthreads();            // the previous point's functions are read
      if (e == 0) {
        zd bj[LB + 1], bh[LB + 1], eiz;
        bessel_jh<LB>(rg * p, bj, bh, eiz);
#pragma unroll
        for (int l = 0; l <= LB; l++) {
          sbj[l] = bj[l];
          sbh[l] = bh[l];
        }
        seiz = times_i(eiz) / rg;
      }
      __syncthreads();
      if (!act) continue;
      const double w = gwwylm[pt];
      const double2 *__restrict__ wy = wylm + pt * nw;
      const zd ez = seiz;
      const zd zlm1 = ml * sbj[l1] - ez * sbh[l1];
      const zd zlm2 = mr * sbj[l2] - ez * sbh[l2];
      const int m3 = m1 - m2, am3 = m3 < 0 ? -m3 : m3;
      const int dl = l1 > l2 ? l1 - l2 : l2 - l1, llow = am3 > dl ? am3 : dl;
      zd zlzl = mk(0.0);
      for (int l3 = l1 + l2; l3 >= llow; l3 -= 2) {
        const double2 wv = wy[(l3 * (l3 + 1)) / 2 + am3];
        zd step = (m3 >= 0) ? mk(wv.x, wv.y) : (double)(1 - 2 * (am3 & 1)) * mk(wv.x, -wv.y);
        if (l3 == 0) {
          if (fabs(rg - rs) < 1e-10) step = step - mk(0.5 * rt4pi);
          else if (rg < rs) step = step - mk(rt4pi);
        }
        zlzl = zlzl + cgnt[l3 / 2 + (lmax_cg + 1) * (lm2 + kkr_cg * lm1)] * step;
      }
      acc = acc + (zlzl * zlm1 * zlm2) * w;
      if (lm1 == lm2 && zj_flag == 1) zj = zj + (zlzl * zlm1 * (ez * sbh[l1] / ml)) * w;
    }
  // outputs: pzz, then the fixed-order sums of pzj and of pzz . tau00
  zd pt = mk(0.0);
  if (act) {
    if (pzz_out) pzz_out[(long)b * kk * kk + lm2 + kk * lm1] = make_double2(acc.x, acc.y);
    const double2 t = tau00[(long)b * kk * kk + lm2 + kk * lm1];
    pt = acc * mk(t.x, t.y);
    if (lm1 == lm2) szj[lm1] = zj;
  }
  red[0][e] = pt.x;
  red[1][e] = pt.y;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (e < s) {
      red[0][e] += red[0][e + s];
      red[1][e] += red[1][e + s];
    }
    __syncthreads();
  }
  if (e == 0) {
    zd pj = mk(0.0);
    if (zj_flag == 1)
      for (int l = 0; l < kk; l++) pj = pj + szj[l];
    if (pzj_out) pzj_out[b] = make_double2(pj.x, pj.y);
    const zd d = mk(red[0][0], red[1][0]) - pj;
    dos_out[b] = make_double2(-scale * d.x / 3.141592653589793, -scale * d.y / 3.141592653589793);
  }
}

__global__ void __launch_bounds__(256) k_interstitial(
    int lmax, int zj_flag, const double2 *__restrict__ prel, const double2 *__restrict__ mleft,
    const double2 *__restrict__ mright, const double *__restrict__ r_sph,
    const int *__restrict__ cell_of, const int *__restrict__ ncrit, int ngaussr, int maxreg,
    const double *__restrict__ grwylm, const double *__restrict__ gwwylm,
    const double2 *__restrict__ wylm, const double *__restrict__ cgnt, int lmax_cg,
    const double2 *__restrict__ tau00, double scale, double2 *__restrict__ pzz_out,
    double2 *__restrict__ pzj_out, double2 *__restrict__ dos_out) {
  const int b = blockIdx.x;
#define LSMSX_INTER(LB)                                                                        \
  interstitial_solve<LB>(b, zj_flag, prel, mleft, mright, r_sph, cell_of, ncrit, ngaussr,      \
                         maxreg, grwylm, gwwylm, wylm, cgnt, lmax_cg, tau00, scale, pzz_out,   \
                         pzj_out, dos_out)
  switch (lmax) {
    case 0: LSMSX_INTER(0); break;
    case 1: LSMSX_INTER(1); break;
    case 2: LSMSX_INTER(2); break;
    default: LSMSX_INTER(3); break;
  }
#undef LSMSX_INTER
}

}  // namespace lsmsx
