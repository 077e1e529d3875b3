// lsmsx_lsms.hpp -- LSMS-shaped host interface of the MI355X tau00 path.
//
// Header-only templates over the reference's own types, so the same calls
// compile inside LSMS (LSMSSystemParameters, LocalTypeInfo, AtomData,
// Matrix<Complex> from src/Main/SystemParameters.hpp, SingleSite/AtomData.hpp,
// include/Matrix.hpp) and against the minimal stand-in types of
// lsmsx_types.hpp used by this package's tests.  Every call goes through the
// C ABI of include/lsmsx.h; nothing here computes on the CPU except the
// kkrsz_ns^2 post-processing that the reference also does on the host.
//
// Fields read (reference declarations):
//   lsms.n_spin_cant, lsms.n_spin_pola, lsms.maxlmax   (SystemParameters.hpp:101-126)
//   local.num_local, local.atom[i], local.tmatStore(r, s), local.blkSizeTmatStore,
//   local.lDimTmatStore                                 (SystemParameters.hpp:209-269)
//   atom.numLIZ, atom.LIZPos(d, i), atom.LIZlmax[i], atom.LIZStoreIdx[i],
//   atom.nrmat, atom.kkrsz, atom.ubr[4], atom.ubrd[4]   (AtomData.hpp:441-467)
//   m(i, j), m.l_dim(), m.resize(n, n)                  (include/Matrix.hpp:126-148)
//   lsms.nrelv, lsms.mtasa, atom.r_mesh, atom.jmt, atom.jws, atom.rInscribed,
//   atom.rws, atom.lmax, vr(j, is)                      (single-site solve)
//
// Error behaviour: like the reference's solvers, a singular pivot (info > 0)
// is printed and the computation continues (zblock_lu_cpp.cpp:58-68), and so
// is an FP64 refinement that did not converge (info -3: the unrefined FP64
// tau00 is used); a refinement wait bound (info -2: tau00 not valid) and a
// device or argument error throw std::runtime_error instead of exit(1).
#pragma once

#include <complex>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lsmsx.h"

namespace lsmsx {

using cplx = std::complex<double>;

// One device context per process, created on first use (the reference keeps
// one DeviceStorage per process, Accelerator/DeviceStorage.hpp:56-184).
struct Runtime {
  lsmsx_ctx *ctx = nullptr;
  int tables_lmax = -1;
  std::mutex mu;
  ~Runtime() {
    if (ctx) lsmsx_ctx_destroy(ctx);
  }
};

inline Runtime &runtime(int device = 0) {
  static Runtime rt;
  std::lock_guard<std::mutex> g(rt.mu);
  if (!rt.ctx) {
    if (lsmsx_ctx_create(device, &rt.ctx) != 0)
      throw std::runtime_error(std::string("lsmsx_ctx_create: ") + lsmsx_last_error());
  }
  return rt;
}

inline void check(int rc, const char *what) {
  if (rc < 0) throw std::runtime_error(std::string(what) + ": " + lsmsx_last_error());
}

// Angular-momentum tables (GauntCoeficients / IFactors) for lsms.maxlmax.
template <class LSMS>
void ensureTables(LSMS &lsms) {
  Runtime &rt = runtime();
  if (rt.tables_lmax == lsms.maxlmax) return;
  int lmax = lsms.maxlmax, kkr = (lmax + 1) * (lmax + 1);
  std::vector<double> cg((lmax + 1) * kkr * kkr), il(2 * kkr * kkr), i1(2 * (2 * lmax + 1));
  check(lsmsx_compute_tables(lmax, cg.data(), il.data(), i1.data()), "lsmsx_compute_tables");
  check(lsmsx_set_tables(rt.ctx, lmax, cg.data(), il.data(), i1.data()), "lsmsx_set_tables");
  rt.tables_lmax = lmax;
}

// Counterpart of DeviceStorage::copyTmatStoreToDevice (DeviceStorage_HIP.cpp:282-309):
// call once per energy group after the t-matrix exchange.
template <class Local>
void copyTmatStoreToDevice(Local &local, int groupSize) {
  Runtime &rt = runtime();
  const cplx *p = &local.tmatStore(0, 0);
  int nstore = (int)local.tmatStore.n_col();
  check(lsmsx_upload_tmats(rt.ctx, reinterpret_cast<const double *>(p), (long)local.tmatStore.l_dim(),
                           nstore, (long)local.blkSizeTmatStore, groupSize),
        "lsmsx_upload_tmats");
}

// The same for relativity = full: the kappa-mu store with the group's
// energies (egrd of the energy group), which set the prel^2/E factor of the
// relativistic blocks (setgij.f:39-55).
template <class LSMS, class Local>
void copyTmatStoreToDeviceRelativistic(LSMS &lsms, Local &local, int groupSize,
                                       const cplx *energies) {
  ensureTables(lsms);  // the kappa-mu transform is sized by the tables' lmax
  Runtime &rt = runtime();
  const cplx *p = &local.tmatStore(0, 0);
  int nstore = (int)local.tmatStore.n_col();
  check(lsmsx_upload_tmats_rel(rt.ctx, reinterpret_cast<const double *>(p),
                               (long)local.tmatStore.l_dim(), nstore,
                               (long)local.blkSizeTmatStore, groupSize,
                               reinterpret_cast<const double *>(energies)),
        "lsmsx_upload_tmats_rel");
}

// Counterpart of DeviceAtom::copyFromAtom (DeviceStorage_HIP.cpp:380-390) for
// every local atom; the local index is the device atom id.
template <class Local>
void copyLIZToDevice(Local &local) {
  Runtime &rt = runtime();
  for (int i = 0; i < local.num_local; i++) {
    auto &atom = local.atom[i];
    std::vector<double> pos(3 * atom.numLIZ);
    std::vector<int> lm(atom.numLIZ), st(atom.numLIZ);
    for (int s = 0; s < atom.numLIZ; s++) {
      for (int d = 0; d < 3; d++) pos[3 * s + d] = atom.LIZPos(d, s);
      lm[s] = atom.LIZlmax[s];
      st[s] = atom.LIZStoreIdx[s];
    }
    check(lsmsx_set_liz(rt.ctx, i, atom.numLIZ, pos.data(), lm.data(), st.data()), "lsmsx_set_liz");
  }
}

// t0 of the central site, as buildKKRSizeTMatrix (linearSolvers_CPU.cpp:11-42).
template <class LSMS, class Local, class Atom>
std::vector<cplx> centralT(LSMS &lsms, Local &local, Atom &atom, int iie, int ispin) {
  int k = lsms.n_spin_cant * atom.kkrsz;
  long jsm = (lsms.n_spin_pola == lsms.n_spin_cant) ? 0 : (long)k * k * ispin;
  std::vector<cplx> t(k * k);
  for (int j = 0; j < k; j++)
    for (int i = 0; i < k; i++)
      t[i + k * j] = local.tmatStore(i + j * k + iie * local.blkSizeTmatStore + jsm, atom.LIZStoreIdx[0]);
  return t;
}

// buildKKRMatrix replacement (MultipleScattering/buildKKRMatrix.hpp:27-42):
// M = 1 - t G(E) in the reference's site order, built on the device for the
// atom registered as localAtomIndex, copied into m.
template <class LSMS, class Atom, class Mat>
void buildKKRMatrixMI355X(LSMS &lsms, Atom &atom, int localAtomIndex, int ispin, int iie, cplx prel,
                          Mat &m) {
  ensureTables(lsms);
  int n = lsms.n_spin_cant * atom.nrmat;
  m.resize(n, n);
  double pr[2] = {prel.real(), prel.imag()};
  check(lsmsx_build_kkr(runtime().ctx, localAtomIndex, ispin, iie, pr, lsms.n_spin_cant,
                        lsms.n_spin_pola, reinterpret_cast<double *>(&m(0, 0)), (int)m.l_dim()),
        "lsmsx_build_kkr");
}

// Per-solve status of a solve (see the error behaviour above).
inline void reportInfo(const char *where, int atom, int info) {
  if (info == 0) return;
  if (info > 0) {
    std::printf("lsmsx %s: atom %d singular pivot info=%d\n", where, atom, info);
  } else if (info == LSMSX_INFO_REFINE_UNCONVERGED) {
    std::printf("lsmsx %s: atom %d FP64 refinement did not converge; unrefined FP64 tau00 used\n",
                where, atom);
  } else {
    throw std::runtime_error(std::string("lsmsx ") + where + ": atom " + std::to_string(atom) +
                             " solve status " + std::to_string(info) + " (tau00 not valid)");
  }
}

// solveTau00zgetrf replacement (linearSolvers_CPU.cpp:98-123): tau00 =
// [M^-1 [t0;0]]_00 from the caller's KKR matrix m (m is left unchanged).
template <class LSMS, class Local, class Atom, class Mat>
void solveTau00MI355X(LSMS &lsms, Local &local, Atom &atom, int iie, Mat &m, Mat &tau00, int ispin) {
  int n = lsms.n_spin_cant * atom.nrmat, k = lsms.n_spin_cant * atom.kkrsz;
  std::vector<cplx> t0 = centralT(lsms, local, atom, iie, ispin);
  std::vector<cplx> out(k * k);
  int info = 0;
  check(lsmsx_solve_tau00_matrix(runtime().ctx, n, k, reinterpret_cast<const double *>(&m(0, 0)),
                                 (int)m.l_dim(), reinterpret_cast<const double *>(t0.data()),
                                 reinterpret_cast<double *>(out.data()), &info),
        "lsmsx_solve_tau00_matrix");
  reportInfo("solveTau00", -1, info);
  for (int j = 0; j < k; j++)
    for (int i = 0; i < k; i++) tau00(i, j) = out[i + k * j];
}

// tau00 - t0, then the local-frame rotation: calculateTau00MinusT and
// rotateTau00ToLocalFrameNonRelativistic (tau00Postprocess.cpp:52-101; trgtol.f);
// for relativity = full, tau00 itself rotated as dmatp tau00 dmat
// (rotateTau00ToLocalFrameRelativistic, tau00Postprocess.cpp:185-192; tripmt).
template <class LSMS, class Atom>
void postprocessTau00(LSMS &lsms, Atom &atom, const std::vector<cplx> &tau00, const std::vector<cplx> &t0,
                      cplx *tau00_l) {
  int k = lsms.n_spin_cant * atom.kkrsz;
  if ((int)lsms.relativity == 2) {  // Relativity::full
    std::vector<cplx> c((size_t)k * k);
    for (int j = 0; j < k; j++)
      for (int i = 0; i < k; i++) {
        cplx x = 0.0;
        for (int q = 0; q < k; q++) x += atom.dmatp(i, q) * tau00[q + k * j];
        c[i + k * j] = x;
      }
    for (int j = 0; j < k; j++)
      for (int i = 0; i < k; i++) {
        cplx x = 0.0;
        for (int q = 0; q < k; q++) x += c[i + k * q] * atom.dmat(q, j);
        tau00_l[i + k * j] = x;
      }
    return;
  }
  std::vector<cplx> d(k * k);
  for (int i = 0; i < k * k; i++) d[i] = tau00[i] - t0[i];
  if (lsms.n_spin_cant == 2) {
    // ml(:,:,s + 2 s') = sum_{ab} U(s,a) mg_ab UD(b,s'), U = [[u1,u3],[u2,u4]]
    int nr = atom.kkrsz;
    for (int sp = 0; sp < 2; sp++)
      for (int s = 0; s < 2; s++)
        for (int c = 0; c < nr; c++)
          for (int r = 0; r < nr; r++) {
            cplx acc = 0.0;
            for (int b = 0; b < 2; b++)
              for (int a = 0; a < 2; a++)
                acc += atom.ubr[s + 2 * a] * atom.ubrd[b + 2 * sp] * d[(a * nr + r) + (2 * nr) * (b * nr + c)];
            tau00_l[(r + nr * c) + nr * nr * (s + 2 * sp)] = acc;
          }
  } else {
    for (int i = 0; i < k * k; i++) tau00_l[i] = d[i];
  }
}

// calculateTauMatrix replacement for one atom (calculateTauMatrix.cpp:320-837):
// device build + solve, host post-processing into tau00_l.
template <class LSMS, class Local, class Atom>
void calculateTauMatrixMI355X(LSMS &lsms, Local &local, Atom &atom, int localAtomIndex, int ispin,
                              cplx prel, cplx *tau00_l, int iie) {
  ensureTables(lsms);
  int k = lsms.n_spin_cant * atom.kkrsz;
  std::vector<cplx> tau(k * k);
  double pr[2] = {prel.real(), prel.imag()};
  int info = 0;
  check(lsmsx_tau00(runtime().ctx, localAtomIndex, ispin, iie, pr, lsms.n_spin_cant, lsms.n_spin_pola,
                    LSMSX_PREC_FP64, reinterpret_cast<double *>(tau.data()), &info),
        "lsmsx_tau00");
  reportInfo("calculateTauMatrix", localAtomIndex, info);
  postprocessTau00(lsms, atom, tau, centralT(lsms, local, atom, iie, ispin), tau00_l);
}

// calculateAllTauMatrices replacement (calculateTauMatrix.cpp:841-945): all
// local atoms (and the second spin when n_spin_pola != n_spin_cant) of one
// energy in a single batched call instead of the OpenMP loop.
// tau00_l is the reference's Matrix<Complex>(kkrsz_ns^2, num_local*n_spin_pola/n_spin_cant).
template <class LSMS, class Local, class MatL>
void calculateAllTauMatricesMI355X(LSMS &lsms, Local &local, cplx energy, int iie, MatL &tau00_l) {
  ensureTables(lsms);
  const double c2inv = 1.0 / ((2.0 * 137.035999206) * (2.0 * 137.035999206));
  cplx prel = std::sqrt(energy * (1.0 + energy * c2inv));
  int nspin = (lsms.n_spin_pola != lsms.n_spin_cant) ? 2 : 1;
  int nl = local.num_local;
  std::vector<int> ids, spins, slots;
  std::vector<double> pr;
  for (int s = 0; s < nspin; s++)
    for (int i = 0; i < nl; i++) {
      ids.push_back(i);
      spins.push_back(s);
      slots.push_back(iie);
      pr.push_back(prel.real());
      pr.push_back(prel.imag());
    }
  int n = (int)ids.size();
  int k = lsms.n_spin_cant * local.atom[0].kkrsz;
  std::vector<cplx> tau((size_t)n * k * k);
  std::vector<int> info(n);
  check(lsmsx_tau00_batched(runtime().ctx, n, ids.data(), spins.data(), slots.data(), pr.data(),
                            lsms.n_spin_cant, lsms.n_spin_pola, LSMSX_PREC_FP64,
                            reinterpret_cast<double *>(tau.data()), info.data()),
        "lsmsx_tau00_batched");
  for (int b = 0; b < n; b++) {
    reportInfo("calculateAllTauMatrices", ids[b], info[b]);
    auto &atom = local.atom[ids[b]];
    std::vector<cplx> one(tau.begin() + (size_t)b * k * k, tau.begin() + (size_t)(b + 1) * k * k);
    postprocessTau00(lsms, atom, one, centralT(lsms, local, atom, iie, spins[b]),
                     &tau00_l(0, ids[b] + spins[b] * nl));
  }
}

// solveTauFullzgetrf replacement for the Kubo caller (Kubo/CurrentMatrix.cpp:184-190;
// linearSolvers_CPU.cpp:125-144): the build + full solve pair becomes one call,
// tau = M^-1 T (nrmat_ns x nrmat_ns, reference site order) from an explicit
// FP64 inverse of the device-built M.  The reference reads t-matrix slot 0
// (buildNRMatSizeTMatrix ignores its iie argument); iie defaults to 0.
template <class LSMS, class Atom, class Mat>
void calculateTauFullMI355X(LSMS &lsms, Atom &atom, int localAtomIndex, int ispin, cplx prel,
                            Mat &tau, int iie = 0) {
  ensureTables(lsms);
  int n = lsms.n_spin_cant * atom.nrmat;
  tau.resize(n, n);
  double pr[2] = {prel.real(), prel.imag()};
  int info = 0;
  check(lsmsx_tau_full(runtime().ctx, localAtomIndex, ispin, iie, pr, lsms.n_spin_cant,
                       lsms.n_spin_pola, reinterpret_cast<double *>(&tau(0, 0)), (int)tau.l_dim(),
                       &info),
        "lsmsx_tau_full");
  reportInfo("calculateTauFull", -1, info);
}

// Single-site t-matrices of every local atom at the energies of one group,
// written into local.tmatStore as solveSingleScatterers does for each energy
// (Main/solveSingleScatterers.cpp:17-45, the omp loop of
// energyContourIntegration.cpp:426-436): one batched GPU call replaces the
// per-atom, per-spin single_scatterer_nonrel_ calls
// (SingleSite/SingleSiteScattering.cpp:6-54).  Column i, slot iie gets atom
// i's t: the spin blocks one after the other (collinear), or the
// global-frame trltog(t_up, t_dn) for n_spin_cant = 2.  vr[i] is atom i's
// potential Matrix (r_mesh.size() x n_spin_pola); all atoms share one lmax
// and one mesh length.  Returns matom per (energy, atom, spin, l) for callers
// that keep it (solution.matom).
template <class LSMS, class Local, class VR>
std::vector<cplx> solveSingleScatterersMI355X(LSMS &lsms, Local &local, std::vector<VR> &vr,
                                              const std::vector<cplx> &energies) {
  Runtime &rt = runtime();
  const int nat = local.num_local, nsp = lsms.n_spin_pola, ne = (int)energies.size();
  if (nat <= 0 || ne <= 0) return {};
  const int lmax = local.atom[0].lmax, iprpts = (int)local.atom[0].r_mesh.size();
  const int kkr = (lmax + 1) * (lmax + 1), nl = lmax + 1;
  std::vector<double> rv, rm, rsph;
  std::vector<int> jmt, jws;
  for (int i = 0; i < nat; i++) {
    auto &a = local.atom[i];
    if (a.lmax != lmax || (int)a.r_mesh.size() != iprpts)
      throw std::runtime_error("solveSingleScatterersMI355X: atoms differ in lmax or mesh length");
    for (int is = 0; is < nsp; is++) {
      for (int j = 0; j < iprpts; j++) rv.push_back(vr[i](j, is));
      rm.insert(rm.end(), a.r_mesh.begin(), a.r_mesh.end());
      jmt.push_back(a.jmt);
      jws.push_back(a.jws);
      rsph.push_back(lsms.mtasa > 0 ? a.rws : a.rInscribed);
    }
  }
  // prel as solveSingleScatterers forms it (c = 2/alpha in Rydberg units)
  const double c = 2.0 * 137.035999206, c2inv = 1.0 / (c * c);
  std::vector<cplx> prel(ne);
  for (int e = 0; e < ne; e++)
    prel[e] = lsms.nrelv > 0 ? std::sqrt(energies[e])
                             : std::sqrt(energies[e] * (1.0 + energies[e] * c2inv));
  const int npot = nat * nsp;
  std::vector<cplx> matom((size_t)ne * npot * nl), tmat((size_t)ne * npot * kkr * kkr);
  check(lsmsx_single_site_nonrel(rt.ctx, lsms.nrelv, lmax, npot, iprpts, rv.data(), rm.data(),
                                 jmt.data(), jws.data(), rsph.data(), ne,
                                 reinterpret_cast<const double *>(energies.data()),
                                 reinterpret_cast<const double *>(prel.data()),
                                 reinterpret_cast<double *>(matom.data()),
                                 reinterpret_cast<double *>(tmat.data()), nullptr, nullptr),
        "lsmsx_single_site_nonrel");
  const long blk = local.blkSizeTmatStore;
  auto t_of = [&](int e, int i, int is) { return &tmat[(((size_t)e * npot) + i * nsp + is) * kkr * kkr]; };
  for (int i = 0; i < nat; i++)
    for (int e = 0; e < ne; e++) {
      cplx *dst = &local.tmatStore((size_t)e * blk, i);
      if (lsms.n_spin_cant == 2) {  // trltog.f: mg(2kkr x 2kkr) from the two local t
        const cplx *u = local.atom[i].ubr, *ud = local.atom[i].ubrd;
        const cplx *t1 = t_of(e, i, 0), *t2 = t_of(e, i, nsp - 1);
        const int n2 = 2 * kkr;
        for (int q = 0; q < n2 * n2; q++) dst[q] = 0.0;
        for (int ri = 0; ri < 2; ri++)
          for (int ci = 0; ci < 2; ci++) {
            const cplx a = ud[ri] * u[2 * ci], b = ud[2 + ri] * u[2 * ci + 1];
            for (int cc = 0; cc < kkr; cc++)
              for (int rr = 0; rr < kkr; rr++)
                dst[(ri * kkr + rr) + (size_t)(ci * kkr + cc) * n2] =
                    a * t1[rr + (size_t)cc * kkr] + b * t2[rr + (size_t)cc * kkr];
          }
      } else {
        for (int is = 0; is < nsp; is++) {
          const cplx *t = t_of(e, i, is);
          for (int q = 0; q < kkr * kkr; q++) dst[(size_t)is * kkr * kkr + q] = t[q];
        }
      }
    }
  return matom;
}

}  // namespace lsmsx
