"""FP64 factors + FP64 refinement (LSMSX_PREC_FP64_REFINED) and the plain
path's pivot guard -- needs a GPU.

Pivoting inside 64x64 diagonal tiles can meet a small pivot where partial
pivoting over the whole column would not.  On the spin-canted bcc-137
workload (N = 4384) one solve (atom 3, energy slot 19) gets a smallest tile
pivot of 1.3e-3 against LAPACK's 0.68, and the unrefined FP64 tau00 is off by
2.6e-10.  The guard redoes such solves with refinement against the untouched
M; the bar is the plain path's 1e-10 relative (Frobenius) against LAPACK
improved by one step of FP64 refinement.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from lsms_amd import capi
from lsms_amd import configs as CF
from tests import adjoint_ref as R

pytestmark = pytest.mark.gpu
REFINED = capi.PREC_FP64_REFINED


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def lapack_refined_tau00(s, a, ie):
    m = R.oracle_m(s, a, ie)
    k = s.kkrsz_ns
    rhs = np.zeros((m.shape[0], k), dtype=np.complex128)
    rhs[:k] = R.oracle_t0(s, a, ie)
    lu = sla.lu_factor(m)
    x = sla.lu_solve(lu, rhs)
    x = x + sla.lu_solve(lu, rhs - m @ x)
    return x[:k]


def test_refined_mode_c1_against_golden():
    s = CF.config_c1()
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_golden.npz"),
                allow_pickle=False)
    ctx = capi.Context(0)
    ctx.load_system(s)
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    tau, info = ctx.tau00_batched(atoms, iies, [s.prel(i) for i in iies], 16, precision=REFINED)
    iters, _ = ctx.last_stats(len(atoms))
    ctx.close()
    assert np.all(info == 0)
    assert np.all(iters >= 2), iters
    worst = max(rel(tau[n], g["tau00"][atoms[n], iies[n]]) for n in range(len(atoms)))
    assert worst < 1e-10, worst


@pytest.mark.parametrize("nsc,nsp,ispin", [(1, 2, 1), (2, 2, 0)])
def test_refined_mode_spin_variants(nsc, nsp, ispin):
    e, w = CF.contour_semicircle(-0.5, 0.7, 4)
    s = CF.make_system("small", "bcc", 5.42, (1, 1, 1), 6.0, 3, e, w, nspecies=2, seed=5,
                       n_spin_cant=nsc, n_spin_pola=nsp)
    ctx = capi.Context(0)
    ctx.load_system(s)
    k = s.kkrsz_ns
    tau, info = ctx.tau00_batched([0, 1, 0], [0, 2, 3], [s.prel(i) for i in (0, 2, 3)], k,
                                  ispins=[ispin] * 3, precision=REFINED)
    ctx.close()
    assert np.all(info == 0)
    for b, (a, ie) in enumerate([(0, 0), (1, 2), (0, 3)]):
        m = R.oracle_m(s, a, ie, ispin)
        rhs = np.zeros((m.shape[0], k), dtype=np.complex128)
        rhs[:k] = R.oracle_t0(s, a, ie, ispin)
        assert rel(tau[b], np.linalg.solve(m, rhs)[:k]) < 1e-10


def test_pivot_guard_on_small_tile_pivot():
    """The C5 solve with a small tile pivot, next to an ordinary one: the plain
    path and the adjoint route's tau00 meet 1e-10 with the guard (and the
    plain path reports refinement sweeps for the guarded solve)."""
    s = CF.config_c5()
    ctx = capi.Context(0)
    ctx.load_system(s)
    cases = [(3, 19), (0, 3)]
    atoms = [a for a, _ in cases]
    iies = [ie for _, ie in cases]
    pr = [s.prel(i) for i in iies]
    tau, info = ctx.tau00_batched(atoms, iies, pr, 32)
    iters, minpiv = ctx.last_stats(2)
    tau_adj, _, _, info_adj = ctx.tau00_adjoint(atoms, iies, s.energies[iies], 32, lloyd=False)
    ctx.set_pivot_guard(0.0)
    tau_unguarded, _ = ctx.tau00_batched(atoms, iies, pr, 32)
    _, minpiv_tile = ctx.last_stats(2)
    ctx.close()
    assert np.all(info == 0)
    assert minpiv_tile[0] < 1e-2 <= minpiv_tile[1]
    assert iters[0] > 0 and iters[1] == 0
    assert np.array_equal(tau[1], tau_unguarded[1])
    assert np.all(info_adj == 0)
    for b, (a, ie) in enumerate(cases):
        ref = lapack_refined_tau00(s, a, ie)
        assert rel(tau[b], ref) < 1e-10, (a, ie)
        assert rel(tau_adj[b], ref) < 1e-10, (a, ie)


def test_unconverged_refinement_takes_the_column_pivoted_fallback():
    """A refinement held to one sweep cannot converge: the FP64-refined mode
    then solves again by LU with partial pivoting over whole columns of the
    FP64 M (info 0, iters -2), and so does a plain solve that the pivot guard
    sends there (a guard of 1e3 catches every solve); the values agree with
    the plain FP64 ones and with LAPACK at 1e-10."""
    s = CF.config_c1()
    ctx = capi.Context(0)
    ctx.load_system(s)
    atoms = np.array([0, 1, 1])
    iies = np.array([4, 12, 27])
    prels = [s.prel(i) for i in iies]
    plain, info0 = ctx.tau00_batched(atoms, iies, prels, 16)
    ctx.set_refinement(max_sweeps=1, tol=1e-12)
    tau_r, info_r = ctx.tau00_batched(atoms, iies, prels, 16, precision=REFINED)
    it_r, _ = ctx.last_stats(3)
    ctx.set_pivot_guard(1e3)
    tau_g, info_g = ctx.tau00_batched(atoms, iies, prels, 16)
    ctx.set_pivot_guard(1e-2)
    ctx.set_refinement()
    tau_ok, info_ok = ctx.tau00_batched(atoms, iies, prels, 16, precision=REFINED)
    ctx.close()
    assert np.all(info0 == 0) and np.all(info_ok == 0)
    assert np.all(info_r == 0) and np.all(it_r == -2)
    assert np.all(info_g == 0)
    for n in range(3):
        ref = lapack_refined_tau00(s, atoms[n], iies[n])
        assert rel(tau_r[n], ref) < 1e-10
        assert rel(tau_g[n], ref) < 1e-10
        assert rel(tau_ok[n], plain[n]) < 1e-10
