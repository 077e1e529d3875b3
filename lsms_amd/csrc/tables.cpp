// tables.cpp -- angular-momentum tables of the KKR assembly, computed on the
// host from closed forms:
//   cgnt(l3/2, L1, L2) = Int Y_{l1 m1} Y*_{l2 m2} Y_{l3 m3} dOmega, m3 = m2 - m1
//                      = (-1)^m2 sqrt((2l1+1)(2l2+1)(2l3+1)/4pi)
//                        (l1 l2 l3; 0 0 0) (l1 l2 l3; m1 -m2 m3)
//   illp(L, L') = i^(l - l'),  ilp1(l) = i^(l+1)
// The reference obtains the same numbers by Gauss-Legendre quadrature
// (Misc/cgaunt_c.f90:55-101, Misc/gaunt.f; Misc/ifacts_c.f:31-58), storing
// only the l3 = l1+l2, l1+l2-2, ... >= max(|m3|,|l1-l2|) entries and zeroing
// values below 1e-14; the same entries are filled here.
#include <cmath>
#include <cstdlib>

#include "../../include/lsmsx.h"

namespace {

double lfact(int n) { return std::lgamma((double)n + 1.0); }

// Wigner 3j symbol by the Racah formula (log-factorials; l <= 12 here).
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0) return 0.0;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
  double tri = 0.5 * (lfact(j1 + j2 - j3) + lfact(j1 - j2 + j3) + lfact(-j1 + j2 + j3) -
                      lfact(j1 + j2 + j3 + 1));
  double pre = 0.5 * (lfact(j1 + m1) + lfact(j1 - m1) + lfact(j2 + m2) + lfact(j2 - m2) +
                      lfact(j3 + m3) + lfact(j3 - m3));
  int kmin = 0;
  kmin = std::max(kmin, j2 - j3 - m1);
  kmin = std::max(kmin, j1 - j3 + m2);
  int kmax = j1 + j2 - j3;
  kmax = std::min(kmax, j1 - m1);
  kmax = std::min(kmax, j2 + m2);
  double sum = 0.0;
  for (int k = kmin; k <= kmax; k++) {
    double lt = lfact(k) + lfact(j1 + j2 - j3 - k) + lfact(j1 - m1 - k) + lfact(j2 + m2 - k) +
                lfact(j3 - j2 + m1 + k) + lfact(j3 - j1 - m2 + k);
    double term = std::exp(tri + pre - lt);
    sum += (k & 1) ? -term : term;
  }
  int ph = j1 - j2 - m3;
  return (ph & 1) ? -sum : sum;
}

}  // namespace

extern "C" int lsmsx_compute_tables(int lmax, double *cgnt, double *illp, double *ilp1) {
  if (lmax < 0 || lmax > 6) return -2;
  const double pi = 3.141592653589793238462643383279502884;
  int kkr = (lmax + 1) * (lmax + 1);
  for (int i = 0; i < (lmax + 1) * kkr * kkr; i++) cgnt[i] = 0.0;
  for (int j2 = 0; j2 < kkr; j2++) {
    int l2 = (int)std::floor(std::sqrt((double)j2) + 1e-9), m2 = j2 - l2 * (l2 + 1);
    for (int j1 = 0; j1 < kkr; j1++) {
      int l1 = (int)std::floor(std::sqrt((double)j1) + 1e-9), m1 = j1 - l1 * (l1 + 1);
      int m3 = m2 - m1;
      int lo = std::max(std::abs(m3), std::abs(l1 - l2));
      for (int l3 = l1 + l2; l3 >= lo; l3 -= 2) {
        double v = std::sqrt((2.0 * l1 + 1) * (2.0 * l2 + 1) * (2.0 * l3 + 1) / (4.0 * pi)) *
                   wigner3j(l1, l2, l3, 0, 0, 0) * wigner3j(l1, l2, l3, m1, -m2, m3);
        if (m2 & 1) v = -v;
        if (std::fabs(v) < 1.0e-14) v = 0.0;
        cgnt[l3 / 2 + (lmax + 1) * (j1 + kkr * j2)] = v;
      }
    }
  }
  static const double pw[4][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
  for (int a = 0; a < kkr; a++) {
    int la = (int)std::floor(std::sqrt((double)a) + 1e-9);
    for (int b = 0; b < kkr; b++) {
      int lb = (int)std::floor(std::sqrt((double)b) + 1e-9);
      int e = (((la - lb) % 4) + 4) % 4;
      illp[2 * (a + kkr * b)] = pw[e][0];
      illp[2 * (a + kkr * b) + 1] = pw[e][1];
    }
  }
  for (int l = 0; l <= 2 * lmax; l++) {
    int e = (l + 1) % 4;
    ilp1[2 * l] = pw[e][0];
    ilp1[2 * l + 1] = pw[e][1];
  }
  return 0;
}
