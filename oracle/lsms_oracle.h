/*
 * lsms_oracle.h -- CPU restatement of the LSMS multiple-scattering tau00 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load or call it.  The product library (lsms_amd/) never links it and
 * never falls back to it.
 *
 * Conventions (mirroring the reference):
 *   - complex numbers are C99 double _Complex (same layout as
 *     std::complex<double> and numpy complex128);
 *   - matrices are column-major, element (i,j) at a[i + j*ld]
 *     (include/Matrix.hpp:126-148);
 *   - angular-momentum index k = l*(l+1)+m, 0-based (Misc/Indices.hpp:42-58);
 *   - cgnt is an Array3d(lmax+1, kkr, kkr): cgnt(l3/2, j1, j2) at
 *     [l3/2 + (lmax+1)*(j1 + kkr*j2)] (include/Array3d.hpp:111-117).
 */
#ifndef LSMS_ORACLE_H
#define LSMS_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex zc;

/* speed of light in Rydberg units: c = 2/alpha (include/PhysicalConstants.hpp:5-8) */
double or_cphot(void);
/* kappa = sqrt(E (1 + E/c^2)), principal branch (calculateTauMatrix.cpp:846) */
zc or_prel(zc energy);

/* --- tables ------------------------------------------------------------ */
void or_lofk_mofk(int lmax, int *lofk, int *mofk);
void or_gauss_legendre(int n, double x1, double x2, double *x, double *w);
void or_plm_normalized(int lmax, double x, double *plm);
void or_cgaunt(int lmax, double *cgnt);
void or_ifacts(int lmax, zc *illp, zc *ilp1);

/* --- structure constants ----------------------------------------------- */
void or_hankel(zc prel, double r, int lend, zc *hfn);
void or_makegij(int lmaxi, int lmaxj, int lmax_tab, zc prel, const double *rij,
                const double *cgnt, const zc *illp, const zc *ilp1, zc *gij);

/* --- KKR matrix (F77 route of buildKKRMatrix) -------------------------- */
int or_nrmat(int numLIZ, const int *LIZlmax);
int or_build_kkr(int maxlmax, int n_spin_cant, int n_spin_pola, int ispin,
                 int numLIZ, const double *LIZPos, const int *LIZlmax,
                 const int *LIZStoreIdx, const zc *tmatStore, long ldTmatStore,
                 long blkSizeTmatStore, int iie, zc prel, zc *m, int ldm);
void or_t0(int maxlmax, int n_spin_cant, int n_spin_pola, int ispin,
           int storeIdx0, const zc *tmatStore, long ldTmatStore,
           long blkSizeTmatStore, int iie, zc *t0);

/* --- dense complex LU (LAPACK zgetf2 semantics) ------------------------ */
int or_zgetrf(int m, int n, zc *a, int lda, int *ipiv);
int or_zgetrs(int n, int nrhs, const zc *a, int lda, const int *ipiv, zc *b,
              int ldb);
void or_zgemm_nn(int m, int n, int k, zc alpha, const zc *a, int lda,
                 const zc *b, int ldb, zc beta, zc *c, int ldc);

/* --- tau00 solvers ----------------------------------------------------- */
int or_solve_tau00_zgetrf(int nrmat_ns, int kkrsz_ns, zc *m, int ldm,
                          const zc *t0, zc *tau00);
int or_solve_tau00_zblocklu(int nrmat_ns, int kkrsz_ns, int nblk, zc *m,
                            int ldm, const zc *t0, zc *tau00);

/* --- post-processing --------------------------------------------------- */
void or_tau00_minus_t(int kkrsz_ns, const zc *tau00, const zc *t0, zc *out);
void or_spin_trafo(const double *evec, zc *u, zc *ud);
void or_trgtol(int nr, int nc, const zc *u, const zc *ud, const zc *mg,
               zc *ml);

/* --- energy contour ---------------------------------------------------- */
void or_congauss(double ebot, double etop, int ngauss, zc *egrd, zc *dele);

/* --- single-site scattering (single_site_oracle.c) ---------------------- */
/* rwave.f: regular/irregular radial solutions on rad[0..kmax), matom = 1/t_l */
int or_rwave(int srel, int l, zc energy, const double *coef, const double *rad, int jmt,
             int kmax, double r_sph, zc *rr, zc *drr, zc *ri, zc *dri, zc *matom);
/* single_scatterer_nonrel.f through semrel_c.f (solver 2) */
int or_single_scatterer_nonrel(int nrelv, int lmax, zc energy, zc prel, const double *vr,
                               const double *r, int jmt, int jws, double r_sph, int iprpts,
                               zc *matom, zc *tmat, zc *zlr, zc *jlr);
void or_fitpot(const double *r, const double *rv, double *coef, int nfit, int iend);
void or_sph_bessel_jh(int lmax, zc x, zc *bj, zc *hm, zc *eiz);
/* cnewint.f: running integral of f x^ip0 on the mesh x */
void or_cnewint(int nr, const double *x, const zc *f, zc *g, int ip0);
/* muffin-tin / ASA part of green_function.f (gf_local, int_zz_zj, mdosms, mgreen) */
int or_green_mt(int lmax, const zc *tau00, const zc *zlr, const zc *jlr, const double *r_mesh,
                int jmt, int jws, double r_sph, int iprpts, int n_spin_pola, zc *dos, zc *green,
                zc *zz_out, zc *zj_out);
/* spin-canted muffin-tin part of green_function.f with the wx/wy/wz recombination */
int or_green_mt_canted(int lmax, const zc *tau00_l, const zc *zlr, const zc *jlr,
                       const double *r_mesh, int jmt, int jws, double r_sph, int iprpts,
                       const zc *w, zc *dos, zc *green);

void or_interstitial(int zj_flag, int lmax, zc pnrel, const zc *matom_left,
                     const zc *matom_right, double r_sph, int ncrit, int ngaussr,
                     const double *grwylm, const double *gwwylm, const zc *wylm,
                     const double *cgnt, int lmax_cg, zc *pzz, zc *pzj);

/* relativistic KKR blocks (relativity = full) */
void or_clebsch(double *u1, double *u2, int *ind1, int *ind2);
void or_relmtrx(const zc *a, int kkr, zc *b);
int or_build_kkr_rel(int maxlmax, int numLIZ, const double *LIZPos, const int *LIZlmax,
                     const int *LIZStoreIdx, const zc *tmatStore, long ldTmatStore,
                     long blkSizeTmatStore, int iie, zc prel, zc energy, zc *m, int ldm);
void or_tripmt(const zc *u, zc *b, const zc *ust, int n);

#ifdef __cplusplus
}
#endif

#endif
