"""Parity of the HIP path (through the C ABI) with the oracle -- needs a GPU.

Tolerances: the plain path is complex FP64 throughout.  KKR matrices agree
elementwise to 1e-12 of the matrix maximum (different but equivalent
operation order in the Hankel/Legendre recursions); tau00 agrees to 1e-10
relative in the Frobenius norm (the bar of the plain FP64 path), and in
practice to ~1e-13.
"""
import numpy as np
import pytest

from lsms_amd import capi
from lsms_amd import configs as CF
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TAU_RTOL = 1e-10
M_RTOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_tau00(s, a, ie, ispin=0):
    at = s.atoms[a]
    M = O.build_kkr(s.maxlmax, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore,
                    s.blkSizeTmatStore, ie, s.prel(ie), s.n_spin_cant, s.n_spin_pola, ispin)
    t0 = O.t0(s.maxlmax, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, ie,
              s.n_spin_cant, s.n_spin_pola, ispin)
    tau, info = O.solve_tau00_zgetrf(M, t0)
    return tau, M, t0


@pytest.fixture(scope="module")
def c1():
    s = CF.config_c1()
    ctx = capi.Context(0)
    ctx.load_system(s)
    yield s, ctx
    ctx.close()


def test_tables_match_oracle():
    cg, illp, ilp1 = capi.compute_tables(3)
    assert np.abs(cg - O.cgaunt(3)).max() < 1e-14
    il, i1 = O.ifacts(3)
    assert np.array_equal(illp, il) and np.array_equal(ilp1, i1)


@pytest.mark.parametrize("ie", [0, 15, 30])
def test_build_kkr_parity_c1(c1, ie):
    s, ctx = c1
    for a in range(len(s.atoms)):
        at = s.atoms[a]
        Mg = ctx.build_kkr(a, ie, s.prel(ie), s.nrmat_ns(a))
        Mo = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore,
                         s.blkSizeTmatStore, ie, s.prel(ie))
        assert np.abs(Mg - Mo).max() <= M_RTOL * np.abs(Mo).max()


def test_build_kkr_against_golden(c1):
    """Committed fixture made from the reference's own makegij_/setgij_ blocks."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "c1_golden.npz")
    g = np.load(path, allow_pickle=False)
    s, ctx = c1
    for n, ie in enumerate(g["m_energy_index"]):
        Mg = ctx.build_kkr(int(g["m_atom"]), int(ie), s.prel(int(ie)), s.nrmat_ns(0))
        Mr = g["m_ref"][n]
        assert np.abs(Mg - Mr).max() <= M_RTOL * np.abs(Mr).max()


def test_tau00_single_c1(c1):
    s, ctx = c1
    for a, ie in [(0, 0), (1, 12), (0, 30)]:
        tg, info = ctx.tau00(a, ie, s.prel(ie), 16)
        to, _, _ = oracle_tau00(s, a, ie)
        assert info == 0
        assert rel(tg, to) < TAU_RTOL


def test_tau00_batched_c1_all_62(c1):
    """Every (atom, energy) solve of the bcc 15-site workload in one batch,
    against the committed golden tau00 (reference leaves + LAPACK) and the oracle."""
    import os
    s, ctx = c1
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_golden.npz"),
                allow_pickle=False)
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16)
    assert np.all(info == 0)
    worst = 0.0
    for n in range(na * ne):
        ref = g["tau00"][atoms[n], iies[n]]
        worst = max(worst, rel(tau[n], ref))
    assert worst < TAU_RTOL, worst


def test_tau00_chunked_equals_unchunked(c1):
    s, ctx = c1
    atoms = np.array([0, 1, 0, 1, 0, 1, 0])
    iies = np.array([0, 3, 7, 11, 19, 25, 30])
    prels = np.array([s.prel(i) for i in iies])
    t1, _ = ctx.tau00_batched(atoms, iies, prels, 16)
    ctx.set_workspace_limit(2 * 256 * 256 * 16)   # two matrices per chunk
    t2, _ = ctx.tau00_batched(atoms, iies, prels, 16)
    ctx.set_workspace_limit(0)
    assert np.array_equal(t1, t2)


@pytest.mark.parametrize("n,k", [(16, 16), (40, 16), (64, 16), (240, 16), (300, 32), (500, 16),
                                 (130, 64)])
def test_solve_matrix_random(n, k):
    """Caller-built matrices, sizes not multiples of the tile, tiny and k = 32/64."""
    rng = np.random.default_rng(n * 7 + k)
    ctx = capi.Context(0)
    A = np.eye(n) + 0.3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
    t0 = rng.normal(size=(k, k)) + 1j * rng.normal(size=(k, k))
    tau, info = ctx.solve_tau00_matrix(A, t0)
    to, _ = O.solve_tau00_zgetrf(A, t0)
    ctx.close()
    assert info == 0
    assert rel(tau, to) < TAU_RTOL


def test_solve_matrix_identity_problem():
    """inversionTest type 2: G = 0, t = I -> tau00 = I exactly."""
    ctx = capi.Context(0)
    tau, info = ctx.solve_tau00_matrix(np.eye(240, dtype=complex), np.eye(16, dtype=complex))
    ctx.close()
    assert info == 0 and np.array_equal(tau, np.eye(16))


def test_solve_matrix_needs_pivoting():
    """Zero diagonal in the non-central part forces row exchanges."""
    rng = np.random.default_rng(11)
    n, k = 200, 16
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    for i in range(n):
        A[i, i] = 0.0
    t0 = np.eye(k, dtype=complex)
    ctx = capi.Context(0)
    tau, info = ctx.solve_tau00_matrix(A, t0)
    ctx.close()
    to, _ = O.solve_tau00_zgetrf(A, t0)
    assert info == 0
    assert rel(tau, to) < TAU_RTOL


def test_singular_reports_info():
    """A zero column in the non-central part is reported, not hidden."""
    n, k = 128, 16
    A = np.eye(n, dtype=complex)
    A[:, 40] = 0.0
    ctx = capi.Context(0)
    tau, info = ctx.solve_tau00_matrix(A, np.eye(k, dtype=complex))
    ctx.close()
    assert info > 0


def small_system(n_spin_cant=1, n_spin_pola=1, rsteps=None, rcut=6.0, seed=3, ne=5):
    e, w = CF.contour_semicircle(-0.5, 0.7, ne)
    return CF.make_system("small", "bcc", 5.42, (1, 1, 1), rcut, 3, e, w, nspecies=2, seed=seed,
                          n_spin_cant=n_spin_cant, n_spin_pola=n_spin_pola, rsteps=rsteps)


@pytest.mark.parametrize("nsc,nsp,ispin,rsteps", [(1, 2, 0, None), (1, 2, 1, None),
                                                   (2, 2, 0, None), (1, 1, 0, [4.8, 5.5, 9, 9]),
                                                   (2, 2, 0, [4.8, 9, 9, 9])])
def test_spin_and_lmax_variants(nsc, nsp, ispin, rsteps):
    s = small_system(nsc, nsp, rsteps)
    ctx = capi.Context(0)
    ctx.load_system(s)
    kc = s.kkrsz_ns
    for a in range(len(s.atoms)):
        for ie in (0, 2, 4):
            Mg = ctx.build_kkr(a, ie, s.prel(ie), s.nrmat_ns(a), ispin)
            to, Mo, _ = oracle_tau00(s, a, ie, ispin)
            assert np.abs(Mg - Mo).max() <= M_RTOL * np.abs(Mo).max()
            tg, info = ctx.tau00(a, ie, s.prel(ie), kc, ispin)
            assert info == 0
            assert rel(tg, to) < TAU_RTOL
    ctx.close()


def test_single_site_liz():
    """numLIZ = 1: M = I and tau00 = t0."""
    s = small_system(rcut=1.0)
    ctx = capi.Context(0)
    ctx.load_system(s)
    t0 = O.t0(3, int(s.atoms[0].LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, 1)
    tg, info = ctx.tau00(0, 1, s.prel(1), 16)
    ctx.close()
    assert info == 0 and np.abs(tg - t0).max() <= 1e-15 * np.abs(t0).max()


def test_static_limit_prel_zero():
    """E = 0 (prel = 0): the static structure constants of makegij_c.f."""
    s = small_system()
    ctx = capi.Context(0)
    ctx.load_system(s)
    at = s.atoms[0]
    Mg = ctx.build_kkr(0, 2, 0.0, s.nrmat_ns(0))
    Mo = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore, s.blkSizeTmatStore,
                     2, 0.0)
    ctx.close()
    assert np.abs(Mg - Mo).max() <= M_RTOL * np.abs(Mo).max()


def test_uneven_batch():
    """LIZ sizes differ inside one batch (bcc, two cut radii)."""
    e, w = CF.contour_semicircle(-0.5, 0.7, 4)
    s = CF.make_system("uneven", "bcc", 5.42, (2, 2, 2), 6.0, 3, e, w, nspecies=2, seed=9,
                       local_atoms=[0, 1, 2, 3], rcut_per_atom=[6.0, 8.0, 5.0, 7.7])
    sizes = sorted({a.numLIZ for a in s.atoms})
    assert len(sizes) >= 3
    ctx = capi.Context(0)
    ctx.load_system(s)
    atoms = np.repeat(np.arange(4), 4)
    iies = np.tile(np.arange(4), 4)
    prels = np.array([s.prel(i) for i in iies])
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16)
    ctx.close()
    assert np.all(info == 0)
    for n in range(len(atoms)):
        to, _, _ = oracle_tau00(s, atoms[n], iies[n])
        assert rel(tau[n], to) < TAU_RTOL


def test_full_size_c2_against_lapack():
    """fcc 135-site (N = 2160) at two contour points, against LAPACK zgetrf on
    the oracle's matrix (the oracle's own LU is too slow at this size)."""
    import scipy.linalg as sl
    s = CF.config_c2(natoms=2)
    ctx = capi.Context(0)
    ctx.load_system(s)
    for a, ie in [(0, 0), (1, 15)]:
        at = s.atoms[a]
        assert s.nrmat_ns(a) == 2160
        M = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore,
                        s.blkSizeTmatStore, ie, s.prel(ie))
        t0 = O.t0(3, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, ie)
        lu, piv, _ = sl.lapack.zgetrf(M)
        rhs = np.zeros((M.shape[0], 16), complex)
        rhs[:16] = t0
        x, _ = sl.lapack.zgetrs(lu, piv, rhs)
        tg, info = ctx.tau00(a, ie, s.prel(ie), 16)
        assert info == 0
        assert rel(tg, x[:16]) < TAU_RTOL
    ctx.close()
