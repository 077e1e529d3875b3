"""Mixed-precision tau00 (FP32 factors on MFMA + FP64 iterative refinement
against the FP64 KKR matrix) -- needs a GPU.

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


def test_mixed_c1_all_62_against_golden():
    s = CF.config_c1()
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_golden.npz"),
                allow_pickle=False)
    ctx = capi.Context(0)
    ctx.load_system(s)
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16, precision=MIXED)
    iters, _ = ctx.last_stats(len(atoms))
    ctx.close()
    assert np.all(info == 0)
    assert np.all(iters >= 2), iters
    worst = max(rel(tau[n], g["tau00"][atoms[n], iies[n]]) for n in range(len(atoms)))
    assert worst < TAU_RTOL, worst


def test_mixed_fcc135_near_real_axis():
    """N = 2160 at contour points and at Im E = 0.001-0.01 Ry (the C3 energies)."""
    s = CF.config_c3(natoms=2)
    ctx = capi.Context(0)
    ctx.load_system(s)
    sel = [(0, 0), (1, 16), (0, 31), (1, 32), (0, 40), (1, 47)]
    atoms = np.array([a for a, _ in sel])
    iies = np.array([e for _, e in sel])
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16, precision=MIXED)
    iters, _ = ctx.last_stats(len(sel))
    ctx.close()
    assert np.all(info == 0)
    for n, (a, ie) in enumerate(sel):
        err = rel(tau[n], lapack_tau00(s, a, ie))
        assert err < TAU_RTOL, (a, ie, err, iters[n])


def test_mixed_equals_plain_within_tolerance():
    s = CF.config_c1()
    ctx = capi.Context(0)
    ctx.load_system(s)
    atoms = np.array([0, 1, 0, 1])
    iies = np.array([0, 7, 19, 30])
    prels = np.array([s.prel(i) for i in iies])
    t64, _ = ctx.tau00_batched(atoms, iies, prels, 16)
    t32, _ = ctx.tau00_batched(atoms, iies, prels, 16, precision=MIXED)
    ctx.close()
    for n in range(4):
        assert rel(t32[n], t64[n]) < TAU_RTOL


def test_mixed_fallback_to_fp64():
    """With one refinement sweep allowed nothing converges: every solve is
    redone by the plain FP64 path and still meets the bar."""
    s = CF.config_c1()
    ctx = capi.Context(0)
    ctx.load_system(s)
    ctx.set_refinement(max_sweeps=1, tol=1e-12)
    atoms = np.array([0, 1, 1])
    iies = np.array([2, 9, 28])
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16, precision=MIXED)
    iters, _ = ctx.last_stats(3)
    ctx.close()
    assert np.all(iters == -1)
    for n in range(3):
        assert rel(tau[n], lapack_tau00(s, atoms[n], iies[n])) < TAU_RTOL


@pytest.mark.parametrize("nsc,nsp,ispin", [(1, 2, 1), (2, 2, 0)])
def test_mixed_spin_variants(nsc, nsp, ispin):
    e, w = CF.contour_semicircle(-0.5, 0.7, 4)
    s = CF.make_system("sp", "bcc", 5.42, (1, 1, 1), 8.0, 3, e, w, seed=12,
                       n_spin_cant=nsc, n_spin_pola=nsp)
    ctx = capi.Context(0)
    ctx.load_system(s)
    k = s.kkrsz_ns
    atoms = np.array([0, 1, 0])
    iies = np.array([0, 1, 3])
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, k, ispins=np.full(3, ispin), precision=MIXED)
    ctx.close()
    assert np.all(info == 0)
    for n in range(3):
        assert rel(tau[n], lapack_tau00(s, atoms[n], iies[n], ispin)) < TAU_RTOL
