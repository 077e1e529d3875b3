// host_mirror_driver.cpp -- drives the LSMS-shaped C++ interface
// (lsms_amd/host/lsmsx_lsms.hpp) the way calculateTauMatrix.cpp /
// energyContourIntegration.cpp would, on inputs written by
// tests/test_gpu_host_mirror.py, and writes the results for comparison with
// the oracle.
//
// input  (binary, little endian): int32 maxlmax, n_spin_cant, n_spin_pola,
//   num_local, num_store, groupSize, blkSize, lDim, relativity; complex
//   tmatStore[lDim*num_store]; complex energies[groupSize]; per atom: int32
//   numLIZ, double pos[3*numLIZ], int32 lmax[numLIZ], int32 store[numLIZ],
//   complex ubr[4], complex ubrd[4], and for relativity = 2 (full) complex
//   dmat[k*k], dmatp[k*k]
// output: complex tau00_l[groupSize][k*k][num_local*nspin] (calculateAllTauMatricesMI355X),
//   then for atom 0, energy 0: complex M[n*n] (buildKKRMatrixMI355X),
//   complex tau00[k*k] (solveTau00MI355X on that M), complex tau00_l[k*k]
//   (calculateTauMatrixMI355X), and, without relativity, complex tau[n*n]
//   (calculateTauFullMI355X)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../lsms_amd/host/lsmsx_lsms.hpp"
#include "../../lsms_amd/host/lsmsx_types.hpp"

using namespace lsmsx::types;

template <class T>
static void rd(FILE *f, T *p, size_t n) {
  if (fread(p, sizeof(T), n, f) != n) {
    std::fprintf(stderr, "short read\n");
    std::exit(2);
  }
}

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input output\n", argv[0]);
    return 2;
  }
  FILE *fi = std::fopen(argv[1], "rb");
  if (!fi) return 2;
  int hdr[9];
  rd(fi, hdr, 9);
  LSMSSystemParameters lsms;
  lsms.maxlmax = hdr[0];
  lsms.n_spin_cant = hdr[1];
  lsms.n_spin_pola = hdr[2];
  LocalTypeInfo local;
  local.num_local = hdr[3];
  int nstore = hdr[4], group = hdr[5];
  local.blkSizeTmatStore = hdr[6];
  local.lDimTmatStore = hdr[7];
  lsms.relativity = hdr[8];
  local.tmatStore.resize(local.lDimTmatStore, nstore);
  rd(fi, &local.tmatStore(0, 0), (size_t)local.lDimTmatStore * nstore);
  std::vector<Complex> energies(group);
  rd(fi, energies.data(), group);
  local.atom.resize(local.num_local);
  for (auto &a : local.atom) {
    rd(fi, &a.numLIZ, 1);
    a.LIZPos.resize(3, a.numLIZ);
    rd(fi, &a.LIZPos(0, 0), 3 * (size_t)a.numLIZ);
    a.LIZlmax.resize(a.numLIZ);
    a.LIZStoreIdx.resize(a.numLIZ);
    rd(fi, a.LIZlmax.data(), a.numLIZ);
    rd(fi, a.LIZStoreIdx.data(), a.numLIZ);
    rd(fi, a.ubr, 4);
    rd(fi, a.ubrd, 4);
    a.lmax = a.LIZlmax[0];
    a.kkrsz = (a.lmax + 1) * (a.lmax + 1);
    if (lsms.relativity == 2) {
      const int kk = lsms.n_spin_cant * a.kkrsz;
      a.dmat.resize(kk, kk);
      a.dmatp.resize(kk, kk);
      rd(fi, &a.dmat(0, 0), (size_t)kk * kk);
      rd(fi, &a.dmatp(0, 0), (size_t)kk * kk);
    }
    a.nrmat = 0;
    for (int s = 0; s < a.numLIZ; s++) a.nrmat += (a.LIZlmax[s] + 1) * (a.LIZlmax[s] + 1);
  }
  std::fclose(fi);

  // the per-energy-group uploads of energyContourIntegration.cpp:462-475
  if (lsms.relativity == 2)
    lsmsx::copyTmatStoreToDeviceRelativistic(lsms, local, group, energies.data());
  else
    lsmsx::copyTmatStoreToDevice(local, group);
  lsmsx::copyLIZToDevice(local);

  int k = lsms.n_spin_cant * local.atom[0].kkrsz;
  int nspin = (lsms.n_spin_pola != lsms.n_spin_cant) ? 2 : 1;
  FILE *fo = std::fopen(argv[2], "wb");
  for (int ie = 0; ie < group; ie++) {
    Matrix<Complex> tau00_l(k * k, local.num_local * nspin);
    lsmsx::calculateAllTauMatricesMI355X(lsms, local, energies[ie], ie, tau00_l);
    std::fwrite(&tau00_l(0, 0), sizeof(Complex), (size_t)k * k * local.num_local * nspin, fo);
  }
  const double c2inv = 1.0 / ((2.0 * 137.035999206) * (2.0 * 137.035999206));
  Complex prel = std::sqrt(energies[0] * (1.0 + energies[0] * c2inv));
  Matrix<Complex> m;
  lsmsx::buildKKRMatrixMI355X(lsms, local.atom[0], 0, 0, 0, prel, m);
  std::fwrite(&m(0, 0), sizeof(Complex), m.n_row() * m.n_col(), fo);
  Matrix<Complex> tau00(k, k);
  lsmsx::solveTau00MI355X(lsms, local, local.atom[0], 0, m, tau00, 0);
  std::fwrite(&tau00(0, 0), sizeof(Complex), (size_t)k * k, fo);
  std::vector<Complex> tl(k * k);
  lsmsx::calculateTauMatrixMI355X(lsms, local, local.atom[0], 0, 0, prel, tl.data(), 0);
  std::fwrite(tl.data(), sizeof(Complex), (size_t)k * k, fo);
  if (lsms.relativity != 2) {
    Matrix<Complex> tfull;
    lsmsx::calculateTauFullMI355X(lsms, local.atom[0], 0, 0, prel, tfull, 0);
    std::fwrite(&tfull(0, 0), sizeof(Complex), tfull.n_row() * tfull.n_col(), fo);
  }
  std::fclose(fo);
  std::printf("host mirror driver ok: %d atoms x %d energies, k=%d\n", local.num_local, group, k);
  return 0;
}
