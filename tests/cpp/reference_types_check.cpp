// reference_types_check.cpp -- compile-only check that the LSMS-shaped host
// interface (lsms_amd/host/lsmsx_lsms.hpp) instantiates with the reference's
// own types: LSMSSystemParameters and LocalTypeInfo
// (src/Main/SystemParameters.hpp:72-269), AtomData
// (src/SingleSite/AtomData.hpp:441-467) and Matrix<Complex>
// (include/Matrix.hpp).  Built with -fsyntax-only against the reference tree
// by tests/test_native_lib.py; nothing here runs.
#include "Main/SystemParameters.hpp"
#include "../../lsms_amd/host/lsmsx_lsms.hpp"

using lsmsx::cplx;

// the argument types of the reference call sites each template replaces
template void lsmsx::ensureTables<LSMSSystemParameters>(LSMSSystemParameters &);
template void lsmsx::copyTmatStoreToDevice<LocalTypeInfo>(LocalTypeInfo &, int);
template void lsmsx::copyTmatStoreToDeviceRelativistic<LSMSSystemParameters, LocalTypeInfo>(
    LSMSSystemParameters &, LocalTypeInfo &, int, const cplx *);
template void lsmsx::copyLIZToDevice<LocalTypeInfo>(LocalTypeInfo &);
template void lsmsx::buildKKRMatrixMI355X<LSMSSystemParameters, AtomData, Matrix<Complex>>(
    LSMSSystemParameters &, AtomData &, int, int, int, cplx, Matrix<Complex> &);
template void lsmsx::solveTau00MI355X<LSMSSystemParameters, LocalTypeInfo, AtomData,
                                      Matrix<Complex>>(LSMSSystemParameters &, LocalTypeInfo &,
                                                       AtomData &, int, Matrix<Complex> &,
                                                       Matrix<Complex> &, int);
template void lsmsx::calculateTauMatrixMI355X<LSMSSystemParameters, LocalTypeInfo, AtomData>(
    LSMSSystemParameters &, LocalTypeInfo &, AtomData &, int, int, cplx, cplx *, int);
template void lsmsx::calculateAllTauMatricesMI355X<LSMSSystemParameters, LocalTypeInfo,
                                                   Matrix<Complex>>(
    LSMSSystemParameters &, LocalTypeInfo &, cplx, int, Matrix<Complex> &);
template void lsmsx::calculateTauFullMI355X<LSMSSystemParameters, AtomData, Matrix<Complex>>(
    LSMSSystemParameters &, AtomData &, int, int, cplx, Matrix<Complex> &, int);
template std::vector<cplx> lsmsx::solveSingleScatterersMI355X<LSMSSystemParameters,
                                                              LocalTypeInfo, Matrix<Real>>(
    LSMSSystemParameters &, LocalTypeInfo &, std::vector<Matrix<Real>> &,
    const std::vector<cplx> &);
