// lsmsx_types.hpp -- minimal LSMS-shaped types for building and testing the
// host interface (lsmsx_lsms.hpp) without the reference's headers.  Only the
// members that interface reads are declared, with the reference's names and
// layouts: column-major Matrix<T> with data[i + j*lDim] (include/Matrix.hpp),
// LSMSSystemParameters, LocalTypeInfo and AtomData fields on the tau path
// (Main/SystemParameters.hpp:72-164,209-269; SingleSite/AtomData.hpp:441-467).
#pragma once

#include <complex>
#include <cstddef>
#include <vector>

namespace lsmsx {
namespace types {

using Real = double;
using Complex = std::complex<double>;

template <class T>
class Matrix {
 public:
  Matrix() = default;
  Matrix(size_t r, size_t c) { resize(r, c); }
  void resize(size_t r, size_t c) {
    nRow = r;
    nCol = c;
    lDim = r;
    data.assign(r * c, T(0));
  }
  T &operator()(size_t i, size_t j) { return data[i + j * lDim]; }
  const T &operator()(size_t i, size_t j) const { return data[i + j * lDim]; }
  T &operator[](size_t i) { return data[i]; }
  size_t l_dim() const { return lDim; }
  size_t n_row() const { return nRow; }
  size_t n_col() const { return nCol; }

 private:
  size_t nRow = 0, nCol = 0, lDim = 0;
  std::vector<T> data;
};

struct AtomData {
  int numLIZ = 0;
  Matrix<Real> LIZPos;            // 3 x numLIZ, relative to the central site
  std::vector<int> LIZlmax, LIZStoreIdx;
  int nrmat = 0, lmax = 0, kkrsz = 0;
  Complex ubr[4], ubrd[4];        // spin rotation u, u^dagger (spin_trafo.f)
  Matrix<Complex> dmat, dmatp;    // spin rotation of the relativistic (kappa-mu) case
  // radial mesh of the single-site solve (AtomData.hpp: r_mesh, jmt, jws,
  // rInscribed, rws)
  std::vector<Real> r_mesh;
  int jmt = 0, jws = 0;
  Real rInscribed = 0, rws = 0;
};

struct LocalTypeInfo {
  int num_local = 0;
  std::vector<AtomData> atom;
  Matrix<Complex> tmatStore;      // (blkSizeTmatStore * groupSize) x num_store
  int blkSizeTmatStore = 0;
  int lDimTmatStore = 0;
};

struct LSMSSystemParameters {
  int n_spin_cant = 1, n_spin_pola = 1;
  int maxlmax = 3;
  int nrelv = 0, mtasa = 0;       // SystemParameters.hpp: relativity switch, ASA flag
  int relativity = 0;             // Relativity: none = 0, scalar = 1, full = 2
};

}  // namespace types
}  // namespace lsmsx
