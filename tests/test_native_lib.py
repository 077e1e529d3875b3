"""The C-ABI library loads and exports every symbol include/lsmsx.h declares
(CPU only: no compute call that needs a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from lsms_amd import capi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "lsmsx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lsmsx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_expected_entry_points():
    syms = declared_symbols()
    assert set(capi.EXPORTED) == set(syms), (set(capi.EXPORTED) ^ set(syms))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_host_tables_match_oracle():
    for lmax in (1, 2, 3):
        cg, illp, ilp1 = capi.compute_tables(lmax)
        assert np.abs(cg - O.cgaunt(lmax)).max() < 1e-14
        a, b = O.ifacts(lmax)
        assert np.array_equal(illp, a) and np.array_equal(ilp1, b)


def test_null_context_is_an_error_not_a_crash():
    lib = capi.load()
    assert lib.lsmsx_set_profiling(None, 1) < 0
    assert b"null" in lib.lsmsx_last_error()


def test_single_site_and_green_entry_points_reject_a_null_context():
    """The single-site, store-fill and Green-function entry points return an
    error for a null context before touching the device (no GPU needed)."""
    lib = capi.load()
    calls = [
        lambda: lib.lsmsx_single_site_nonrel(None, 0, 3, 1, 10, None, None, None, None, None, 1,
                                             None, None, None, None, None, None),
        lambda: lib.lsmsx_green_mt(None, 1, 3, 10, 1, 1, None, None, None, None, None, None, None,
                                   None, None, None),
        lambda: lib.lsmsx_green_mt_canted(None, 1, 3, 10, 1, None, None, None, None, None, None,
                                          None, None, None, None, None),
        lambda: lib.lsmsx_green_mt_resident(None, 1, 3, 10, 1, None, None, None, None, None),
        lambda: lib.lsmsx_set_single_site_resident(None, 1),
        lambda: lib.lsmsx_tmats_from_single_site(None, 1, 1, 1, 1, None, None),
        lambda: lib.lsmsx_tmats_from_single_site_canted(None, 1, 1, 1, 1, 1, None, None, None,
                                                        None, None),
    ]
    for call in calls:
        assert call() < 0
        assert b"null" in lib.lsmsx_last_error()


REF_TREE = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF_TREE, "src", "Main")),
                    reason="reference tree not present")
def test_host_interface_compiles_against_reference_types():
    """lsms_amd/host/lsmsx_lsms.hpp instantiates with the reference's own
    LSMSSystemParameters, LocalTypeInfo, AtomData and Matrix<Complex>
    (tests/cpp/reference_types_check.cpp, compile only: the headers need the
    image's MPI headers, nothing is linked or run)."""
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", "reference_types_check.cpp")
    inc = [f"-I{REF_TREE}/include", f"-I{REF_TREE}/src"]
    inc += [f"-I{d}" for d in ("/opt/conda/include",) if os.path.exists(os.path.join(d, "mpi.h"))]
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", *inc, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
