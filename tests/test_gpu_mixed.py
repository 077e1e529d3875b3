"""Mixed-precision tau00 (FP32 factors on MFMA + FP64 iterative refinement
against the FP64 KKR matrix) -- needs a GPU.  Both mixed modes are covered:
FP32 trailing updates, and split-FP16 trailing updates (hi/lo FP16 panels on
the FP16 matrix cores, FP32 accumulation).

Bar: tau00 within 1e-10 relative (Frobenius) of the FP64 oracle / LAPACK,
including energies 0.001 Ry above the real axis; solves whose refinement
stalls are redone in FP64 (forced here by allowing a single sweep).
"""
import os

import numpy as np
import pytest
import scipy.linalg as sl

from lsms_amd import capi
from lsms_amd import configs as CF
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MIXED = capi.PREC_MIXED_FP32
MODES = [capi.PREC_MIXED_FP32, capi.PREC_MIXED_F16]
TAU_RTOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def lapack_tau00(s, a, ie, ispin=0):
    at = s.atoms[a]
    M = O.build_kkr(s.maxlmax, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore,
                    s.blkSizeTmatStore, ie, s.prel(ie), s.n_spin_cant, s.n_spin_pola, ispin)
    t0 = O.t0(s.maxlmax, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, ie,
              s.n_spin_cant, s.n_spin_pola, ispin)
    lu, piv, _ = sl.lapack.zgetrf(M)
    k = t0.shape[0]
    rhs = np.zeros((M.shape[0], k), complex)
    rhs[:k] = t0
    x, _ = sl.lapack.zgetrs(lu, piv, rhs)
    return x[:k]


@pytest.mark.parametrize("mode", MODES)
def test_mixed_c1_all_62_against_golden(mode):
    s = CF.config_c1()
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_golden.npz"),
                allow_pickle=False)
    ctx = capi.Context(0)
    ctx.load_system(s)
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
    iters, _ = ctx.last_stats(len(atoms))
    ctx.close()
    assert np.all(info == 0)
    assert np.all(iters >= 2), iters
    worst = max(rel(tau[n], g["tau00"][atoms[n], iies[n]]) for n in range(len(atoms)))
    assert worst < TAU_RTOL, worst


@pytest.mark.parametrize("mode", MODES)
def test_mixed_fcc135_near_real_axis(mode):
    """N = 2160 at contour points and at Im E = 0.001-0.01 Ry (the C3 energies)."""
    s = CF.config_c3(natoms=2)
    ctx = capi.Context(0)
    ctx.load_system(s)
    sel = [(0, 0), (1, 16), (0, 31), (1, 32), (0, 40), (1, 47)]
    atoms = np.array([a for a, _ in sel])
    iies = np.array([e for _, e in sel])
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
    iters, _ = ctx.last_stats(len(sel))
    ctx.close()
    assert np.all(info == 0)
    for n, (a, ie) in enumerate(sel):
        err = rel(tau[n], lapack_tau00(s, a, ie))
        assert err < TAU_RTOL, (a, ie, err, iters[n])


@pytest.mark.parametrize("mode", MODES)
def test_mixed_equals_plain_within_tolerance(mode):
    s = CF.config_c1()
    ctx = capi.Context(0)
    ctx.load_system(s)
    atoms = np.array([0, 1, 0, 1])
    iies = np.array([0, 7, 19, 30])
    prels = np.array([s.prel(i) for i in iies])
    t64, _ = ctx.tau00_batched(atoms, iies, prels, 16)
    t32, _ = ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
    ctx.close()
    for n in range(4):
        assert rel(t32[n], t64[n]) < TAU_RTOL


@pytest.mark.parametrize("mode", MODES)
def test_mixed_fallback_to_fp64(mode):
    """With one refinement sweep allowed nothing converges, the FP64-refined
    redo cannot either, and every solve is solved again by the column-pivoted
    fallback (LU over whole columns of the FP64 M): info 0, iters -2, and the
    bar against LAPACK."""
    s = CF.config_c1()
    ctx = capi.Context(0)
    ctx.load_system(s)
    ctx.set_refinement(max_sweeps=1, tol=1e-12)
    atoms = np.array([0, 1, 1])
    iies = np.array([2, 9, 28])
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
    iters, _ = ctx.last_stats(3)
    ctx.close()
    assert np.all(iters == -2)
    assert np.all(info == 0), info
    for n in range(3):
        assert rel(tau[n], lapack_tau00(s, atoms[n], iies[n])) < TAU_RTOL


@pytest.mark.parametrize("mode", MODES)
def test_mixed_chunked_equals_unchunked(mode):
    """A batch split into chunks of one solve (one stream, no halves) and of a
    few solves gives the unchunked results and sweep counts bitwise; an empty
    batch returns at once."""
    s = CF.config_c1()
    ctx = capi.Context(0)
    ctx.load_system(s)
    atoms = np.array([0, 1, 0, 1, 0, 1, 0])
    iies = np.array([0, 3, 7, 11, 19, 25, 30])
    prels = np.array([s.prel(i) for i in iies])
    t1, i1 = ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
    it1, _ = ctx.last_stats(len(atoms))
    for limit in (1, 12_000_000):
        ctx.set_workspace_limit(limit)
        t2, i2 = ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
        it2, _ = ctx.last_stats(len(atoms))
        assert np.array_equal(t1, t2) and np.array_equal(i1, i2), limit
        assert np.array_equal(it1, it2), limit
    ctx.set_workspace_limit(0)
    t0, i0 = ctx.tau00_batched(np.zeros(0, int), np.zeros(0, int), np.zeros(0, complex), 16,
                               precision=mode)
    ctx.close()
    assert t0.shape == (0, 16, 16) and i0.shape == (0,)
    assert np.all(i1 == 0) and np.all(it1 >= 2)


@pytest.mark.parametrize("mode", MODES)
def test_mixed_tile_counts_one_to_seven(mode):
    """Solves of 1, 2, 5 and 7 tiles (N = 16, 112, 304, 432) in one batch:
    the one-launch refinement sweep deals tile rows to eight waves, so this
    covers waves with no tile, one tile, and rows split unevenly."""
    e, w = CF.contour_semicircle(-0.5, 0.7, 4)
    s = CF.make_system("tiles", "sc", 5.0, (2, 2, 2), 0, 3, e, w, seed=3,
                       local_atoms=[0, 1, 2, 3], rcut_per_atom=[3.0, 5.5, 7.5, 9.0])
    assert [s.nrmat_ns(i) for i in range(4)] == [16, 112, 304, 432]
    ctx = capi.Context(0)
    ctx.load_system(s)
    atoms = np.array([0, 1, 2, 3, 3, 0])
    iies = np.array([0, 1, 2, 3, 0, 3])
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
    iters, _ = ctx.last_stats(len(atoms))
    ctx.close()
    assert np.all(info == 0)
    assert np.all(iters >= 2), iters
    for n in range(len(atoms)):
        assert rel(tau[n], lapack_tau00(s, atoms[n], iies[n])) < TAU_RTOL


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("nsc,nsp,ispin", [(1, 2, 1), (2, 2, 0)])
def test_mixed_spin_variants(nsc, nsp, ispin, mode):
    e, w = CF.contour_semicircle(-0.5, 0.7, 4)
    s = CF.make_system("sp", "bcc", 5.42, (1, 1, 1), 8.0, 3, e, w, seed=12,
                       n_spin_cant=nsc, n_spin_pola=nsp)
    ctx = capi.Context(0)
    ctx.load_system(s)
    k = s.kkrsz_ns
    atoms = np.array([0, 1, 0])
    iies = np.array([0, 1, 3])
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, k, ispins=np.full(3, ispin), precision=mode)
    ctx.close()
    assert np.all(info == 0)
    for n in range(3):
        assert rel(tau[n], lapack_tau00(s, atoms[n], iies[n], ispin)) < TAU_RTOL


def t0_blocks(s, a, iies):
    at = s.atoms[a]
    return np.array([O.t0(s.maxlmax, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore,
                          int(ie)) for ie in iies])


@pytest.mark.parametrize("mode", MODES)
def test_c3_per_energy_error_and_idos(mode):
    """The 48-energy workload (32 contour points + 16 at Im E = 0.001-0.01 Ry):
    per-energy mixed-vs-FP64 tau00 error within 1e-10, the FP64 path against
    LAPACK at the nearest-axis energies, and the multiple-scattering IDOS
    (unit radial integrals) of the mixed path within 1e-10 relative of FP64."""
    from lsms_amd import contour
    s = CF.config_c3(natoms=4)
    ctx = capi.Context(0)
    ctx.load_system(s)
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    prels = np.array([s.prel(i) for i in iies])
    t32, info32 = ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
    t64, info64 = ctx.tau00_batched(atoms, iies, prels, 16)
    ctx.close()
    assert np.all(info32 == 0) and np.all(info64 == 0)
    err = np.array([rel(t32[n], t64[n]) for n in range(len(atoms))]).reshape(na, ne)
    per_e = err.max(axis=0)
    assert per_e.max() < TAU_RTOL, per_e
    near = np.argsort(np.imag(s.energies))[:2]  # Im E = 0.001, ...
    for ie in near:
        assert rel(t64[ie], lapack_tau00(s, 0, ie)) < TAU_RTOL
    w = s.weights
    for a in range(na):
        sl_ = slice(a * ne, (a + 1) * ne)
        t0 = t0_blocks(s, a, range(ne))
        n64 = contour.idos_ms(t64[sl_], t0, w)
        n32 = contour.idos_ms(t32[sl_], t0, w)
        assert abs(n32 - n64) < 1e-10 * abs(n64), (a, n32, n64)


def test_lloyd_idos_c1_against_oracle():
    """Lloyd IDOS of both C1 atoms over the 31-point contour: GPU adjoint route
    against the oracle's four-point-derivative traces."""
    from lsms_amd import contour
    from tests import adjoint_ref as R
    s = CF.config_c1()
    ctx = capi.Context(0)
    ctx.load_system(s)
    ctx.upload_dtmats(s.dtmat_store(), s.blkSizeTmatStore, len(s.energies))
    ne = len(s.energies)
    for a in range(len(s.atoms)):
        _, ll, _, info = ctx.tau00_adjoint([a] * ne, np.arange(ne), s.energies, 16)
        assert np.all(info == 0)
        ref = [R.lloyd_trace(s, a, e) for e in s.energies]
        n_gpu = contour.idos_lloyd(ll, s.weights)
        n_ref = contour.idos_lloyd(ref, s.weights)
        assert abs(n_gpu - n_ref) < 1e-9 * max(1.0, abs(n_ref)), (a, n_gpu, n_ref)
    ctx.close()


@pytest.mark.parametrize("mode", MODES)
def test_residual_forms_fcc135(mode):
    """The residual from the structure constants stored once per site pair
    (B - X + t G X, the default for these batches) and from the FP64 M both
    refine fcc-135 solves (N = 2160, contour and Im E = 0.001 Ry) to the
    bar against LAPACK, in about the same number of sweeps."""
    s = CF.config_c3(natoms=2)
    ctx = capi.Context(0)
    ctx.load_system(s)
    sel = [(0, 3), (1, 20), (0, 32), (1, 47)]
    atoms = np.array([a for a, _ in sel])
    iies = np.array([e for _, e in sel])
    prels = np.array([s.prel(i) for i in iies])
    out = {}
    for form in (capi.RESIDUAL_AUTO, capi.RESIDUAL_MATRIX):
        ctx.set_residual_form(form)
        tau, info = ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
        iters, _ = ctx.last_stats(len(sel))
        assert np.all(info == 0) and np.all(iters > 0), (form, info, iters)
        out[form] = (tau, iters)
    ctx.close()
    for n, (a, ie) in enumerate(sel):
        ref = lapack_tau00(s, a, ie)
        for form, (tau, iters) in out.items():
            err = rel(tau[n], ref)
            print(f"form {form} mode {mode} solve {a},{ie}: {err:.2e} sweeps {iters[n]}")
            assert err < TAU_RTOL, (form, a, ie, err)
    assert abs(int(out[0][1].sum()) - int(out[1][1].sum())) <= len(sel)
