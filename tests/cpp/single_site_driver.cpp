// single_site_driver.cpp -- drives solveSingleScatterersMI355X of the
// LSMS-shaped C++ interface (lsms_amd/host/lsmsx_lsms.hpp) the way
// energyContourIntegration.cpp would, on inputs written by
// tests/test_gpu_host_mirror.py, and writes the filled t-matrix store.
//
// input  (binary, little endian): int32 nrelv, mtasa, n_spin_cant,
//   n_spin_pola, num_local, lmax, iprpts, groupSize, blkSize, lDim;
//   complex energies[groupSize]; per atom: double r_mesh[iprpts], int32 jmt,
//   jws, double rInscribed, rws, complex ubr[4], ubrd[4],
//   double vr[iprpts * n_spin_pola] (column per spin)
// output: complex tmatStore[lDim * num_local]
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
  int hdr[10];
  rd(fi, hdr, 10);
  LSMSSystemParameters lsms;
  lsms.nrelv = hdr[0];
  lsms.mtasa = hdr[1];
  lsms.n_spin_cant = hdr[2];
  lsms.n_spin_pola = hdr[3];
  LocalTypeInfo local;
  local.num_local = hdr[4];
  const int lmax = hdr[5], iprpts = hdr[6], group = hdr[7];
  local.blkSizeTmatStore = hdr[8];
  local.lDimTmatStore = hdr[9];
  local.tmatStore.resize(local.lDimTmatStore, local.num_local);
  std::vector<Complex> energies(group);
  rd(fi, energies.data(), group);
  local.atom.resize(local.num_local);
  std::vector<Matrix<Real>> vr(local.num_local);
  for (int i = 0; i < local.num_local; i++) {
    AtomData &a = local.atom[i];
    a.lmax = lmax;
    a.kkrsz = (lmax + 1) * (lmax + 1);
    a.r_mesh.resize(iprpts);
    rd(fi, a.r_mesh.data(), iprpts);
    rd(fi, &a.jmt, 1);
    rd(fi, &a.jws, 1);
    rd(fi, &a.rInscribed, 1);
    rd(fi, &a.rws, 1);
    rd(fi, a.ubr, 4);
    rd(fi, a.ubrd, 4);
    vr[i].resize(iprpts, lsms.n_spin_pola);
    rd(fi, &vr[i](0, 0), (size_t)iprpts * lsms.n_spin_pola);
  }
  std::fclose(fi);
  lsmsx::solveSingleScatterersMI355X(lsms, local, vr, energies);
  FILE *fo = std::fopen(argv[2], "wb");
  std::fwrite(&local.tmatStore(0, 0), sizeof(Complex),
              (size_t)local.lDimTmatStore * local.num_local, fo);
  std::fclose(fo);
  std::printf("single-site driver ok: %d atoms x %d energies\n", local.num_local, group);
  return 0;
}
