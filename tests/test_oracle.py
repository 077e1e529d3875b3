"""Pins of the CPU oracle (oracle/lsms_oracle.c) -- CPU only.

Each oracle function is checked against something that does not share its
formula: golden values from the reference's own tests, the reference's leaf
routines compiled from source (oracle/_ref, when present), closed forms and
independent library routines (scipy special functions, LAPACK), and
mathematical invariants of the tau00 problem.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.special as sp

from oracle import oracle as O
from oracle import ref_leaf as R
from lsms_amd import configs as CF

needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (needs the reference tree)")


# ---------------------------------------------------------------- Gaunt numbers
def test_gaunt_golden_values():
    """Test/Unit/test_gaunt.cpp:84-87 (tolerance 1e-12 as there)."""
    c = O.cgaunt(3)
    assert abs(c[0, 0, 0] - 0.28209479177387842) < 1e-12
    assert abs(c[1, 6, 6] - 0.1802237515728685748747963944546917104709985754562351930533759359) < 1e-12


def _w3j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j by the Racah sum in exact integers (fractions), then sqrt."""
    from fractions import Fraction
    if m1 + m2 + m3 != 0 or j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    f = math.factorial
    tri = Fraction(f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3), f(j1 + j2 + j3 + 1))
    pre = f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3)
    s = Fraction(0)
    for k in range(0, j1 + j2 - j3 + 1):
        args = [k, j1 + j2 - j3 - k, j1 - m1 - k, j2 + m2 - k, j3 - j2 + m1 + k, j3 - j1 - m2 + k]
        if min(args) < 0:
            continue
        den = 1
        for a in args:
            den *= f(a)
        s += Fraction((-1) ** k, den)
    val2 = tri * pre * s * s  # square of the magnitude, exact
    sign = (-1) ** (j1 - j2 - m3) * (1 if s >= 0 else -1)
    return sign * math.sqrt(float(val2))


@pytest.mark.parametrize("lmax", [1, 2, 3])
def test_gaunt_closed_form(lmax):
    """cgnt(l3/2,j1,j2) = Int Y_{l1m1} Y*_{l2m2} Y_{l3,m2-m1}: closed form
    (-1)^m2 sqrt((2l1+1)(2l2+1)(2l3+1)/4pi) (l1 l2 l3;000)(l1 l2 l3;m1 -m2 m3)."""
    c = O.cgaunt(lmax)
    kkr = (lmax + 1) ** 2
    worst = 0.0
    for j2 in range(kkr):
        l2 = int(math.isqrt(j2)); m2 = j2 - l2 * (l2 + 1)
        for j1 in range(kkr):
            l1 = int(math.isqrt(j1)); m1 = j1 - l1 * (l1 + 1)
            m3 = m2 - m1
            for l3 in range(l1 + l2, max(abs(m3), abs(l1 - l2)) - 1, -2):
                v = ((-1) ** m2 * math.sqrt((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / (4 * math.pi))
                     * _w3j(l1, l2, l3, 0, 0, 0) * _w3j(l1, l2, l3, m1, -m2, m3))
                worst = max(worst, abs(c[l3 // 2, j1, j2] - v))
    assert worst < 1e-14


def test_gaunt_selection_rules_zero():
    """Entries outside the stored l3 ladder are zero (parity/triangle/m rules)."""
    c = O.cgaunt(3)
    # l1 = l2 = 0 has only l3 = 0
    assert np.all(c[1:, 0, 0] == 0.0)


@needs_ref
@pytest.mark.parametrize("lmax", [1, 2, 3, 4])
def test_gaunt_vs_reference(lmax):
    assert np.array_equal(O.cgaunt(lmax), R.cgaunt(lmax))


@needs_ref
def test_ifacts_vs_reference():
    a, b = O.ifacts(3)
    c, d = R.ifacts(3)
    assert np.abs(a - c).max() == 0.0 and np.abs(b - d).max() == 0.0


def test_ifacts_values():
    illp, ilp1 = O.ifacts(3)
    for lm in range(16):
        for lmp in range(16):
            assert illp[lm, lmp] == 1j ** (math.isqrt(lm) - math.isqrt(lmp))
    assert np.allclose(ilp1, [1j ** (l + 1) for l in range(7)])


# ---------------------------------------------------------- special functions
@pytest.mark.parametrize("x", [-0.97, -0.3, 0.0, 0.41, 0.999])
def test_plm_vs_scipy_ylm(x):
    """Y_lm(theta,0) = Pbar_lm(cos theta), Condon-Shortley phase included."""
    lmax = 6
    p = O.plm_normalized(lmax, x)
    th = math.acos(x)
    for l in range(lmax + 1):
        for m in range(l + 1):
            y = sp.sph_harm_y(l, m, th, 0.0)
            assert abs(p[l * (l + 1) // 2 + m] - y.real) < 1e-13


@needs_ref
def test_plm_gl_vs_reference():
    for x in np.linspace(-0.99, 0.99, 7):
        assert np.array_equal(O.plm_normalized(6, x), R.plm_normalized(6, x))
    for n in (7, 14, 31):
        a, b = O.gauss_legendre(n)
        c, d = R.gauss_legendre_points(n)
        assert np.array_equal(a, c) and np.array_equal(b, d)


def test_gauss_legendre_vs_numpy():
    for n in (7, 14, 31):
        x, w = O.gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - xr).max() < 1e-14 and np.abs(w - wr).max() < 1e-14


def _h1_series(l, z):
    """Spherical Hankel h_l^(1)(z) from its finite closed form
    h_l(z) = (-i)^(l+1) e^{iz}/z sum_k i^k (l+k)! / (k! (l-k)! (2z)^k)."""
    s = 0j
    for k in range(l + 1):
        s += (1j ** k) * math.factorial(l + k) / (math.factorial(k) * math.factorial(l - k) * (2 * z) ** k)
    return ((-1j) ** (l + 1)) * np.exp(1j * z) / z * s


@pytest.mark.parametrize("e", [-0.5 + 0.01j, 0.3 + 0.4j, 0.7 + 0.003j, -0.2 + 0.1j])
def test_hankel_closed_form(e):
    """hfn_l = -i^(l+1) kappa h_l(kappa r)."""
    kap = O.prel(e)
    for r in (2.7, 4.69, 9.4):
        h = O.hankel(kap, r, 6)
        for l in range(7):
            ref = -(1j ** (l + 1)) * kap * _h1_series(l, kap * r)
            assert abs(h[l] - ref) <= 1e-12 * abs(ref)


def test_hankel_vs_scipy_bessel():
    """Where j_l + i y_l does not cancel (real-ish kappa), scipy agrees too."""
    kap = O.prel(0.3 + 0.05j)
    for r in (2.7, 4.69):
        h = O.hankel(kap, r, 6)
        for l in range(7):
            z = kap * r
            ref = -(1j ** (l + 1)) * kap * (sp.spherical_jn(l, z) + 1j * sp.spherical_yn(l, z))
            assert abs(h[l] - ref) <= 1e-10 * abs(ref)


def test_hankel_static_limit():
    """prel = 0: hfn_l = i^l (2l-1)!! / r^(l+1) (makegij_c.f static branch)."""
    r = 3.3
    h = O.hankel(0.0, r, 6)
    for l in range(7):
        df = 1
        for k in range(1, 2 * l, 2):
            df *= k
        assert abs(h[l] - (1j ** l) * df / r ** (l + 1)) < 1e-15 * abs(h[l]) + 1e-300


def test_prel():
    c = 2 * 137.035999206
    for e in (-0.5 + 0.01j, 0.7 + 0.003j):
        assert abs(O.prel(e) - np.sqrt(e * (1 + e / c ** 2))) < 1e-15


# ------------------------------------------------------------ structure constants
@needs_ref
def test_makegij_vs_reference():
    rng = np.random.default_rng(7)
    cg = O.cgaunt(3)
    il, i1 = O.ifacts(3)
    for t in range(300):
        rij = rng.normal(size=3) * 4.0
        if t % 10 == 0:
            rij[:2] = 0.0          # z axis: rho < 1e-6 branch
        if t % 10 == 1:
            rij[:2] = 1e-8
        e = complex(rng.uniform(-0.6, 0.8), rng.uniform(0.0, 0.6))
        p = 0.0 if t % 25 == 3 else O.prel(e)
        li, lj = (int(v) for v in rng.integers(0, 4, 2))
        a = O.makegij(li, lj, 3, p, rij, cg, il, i1)
        b = R.makegij(li, lj, 3, p, rij)
        assert np.array_equal(a, b), (t, li, lj)


def test_makegij_reciprocity():
    """Parity of the structure constants: G(-R)(L2,L1) = (-1)^(l1+l2) G(R)(L2,L1),
    since Y_lm(-R) = (-1)^l Y_lm(R) and only l3 = l1+l2 (mod 2) contribute."""
    rij = np.array([1.3, -2.1, 3.4])
    p = O.prel(0.2 + 0.3j)
    g1 = O.makegij(3, 3, 3, p, rij)
    g2 = O.makegij(3, 3, 3, p, -rij)
    ls = np.array([math.isqrt(k) for k in range(16)])
    # G(-R)(L2,L1) = (-1)^(l1+l2) G(R)(L2,L1): parity of the l3 ladder
    ref = g1 * ((-1.0) ** (ls[:, None] + ls[None, :]))
    assert np.abs(g2 - ref).max() < 1e-12 * np.abs(g1).max()


# ------------------------------------------------------------- KKR assembly
def _ref_assemble(sysm, a, ie, ispin=0):
    """M built block by block from the reference's makegij_/setgij_ with the
    t packing of calculateTauMatrix.cpp:155-183 (numpy matmul for zgemm)."""
    at = sysm.atoms[a]
    nsc, nsp = sysm.n_spin_cant, sysm.n_spin_pola
    kkrsz = (sysm.maxlmax + 1) ** 2
    kns = nsc * kkrsz
    p = sysm.prel(ie)
    blk = sysm.blkSizeTmatStore
    kk = [nsc * (l + 1) ** 2 for l in at.LIZlmax]
    off = np.concatenate([[0], np.cumsum(kk)])
    N = off[-1]
    M = np.eye(N, dtype=complex)
    cg = R.cgaunt(sysm.maxlmax)
    il, i1 = R.ifacts(sysm.maxlmax)
    for i in range(at.numLIZ):
        col = sysm.tmatStore[:, at.LIZStoreIdx[i]]
        k1 = (at.LIZlmax[i] + 1) ** 2
        if nsp == nsc:
            T = col[ie * blk:ie * blk + kns * kns].reshape(kns, kns, order="F")
            idx = np.concatenate([np.arange(k1) + s * kkrsz for s in range(nsc)])
            tn = T[np.ix_(idx, idx)]
        else:
            T = col[ie * blk + kkrsz * kkrsz * ispin:ie * blk + kkrsz * kkrsz * (ispin + 1)]
            tn = T.reshape(kkrsz, kkrsz, order="F")[:k1, :k1]
        for j in range(at.numLIZ):
            if i == j:
                continue
            rij = at.LIZPos[:, i] - at.LIZPos[:, j]
            g = R.makegij(int(at.LIZlmax[i]), int(at.LIZlmax[j]), sysm.maxlmax, p, rij, cg, il, i1)
            bg = R.setgij(g, nsc, sysm.energies[ie])
            M[off[i]:off[i + 1], off[j]:off[j + 1]] = -tn @ bg
    return M


@needs_ref
@pytest.mark.parametrize("ie", [0, 15, 30])
def test_build_kkr_vs_reference_blocks(ie):
    s = CF.config_c1()
    at = s.atoms[1]
    M = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore, s.blkSizeTmatStore,
                    ie, s.prel(ie))
    Mr = _ref_assemble(s, 1, ie)
    assert np.abs(M - Mr).max() <= 1e-14 * np.abs(Mr).max()


def small_system(n_spin_cant=1, n_spin_pola=1, rsteps=None, rcut=6.0, lmax=3, seed=3):
    e, w = CF.contour_semicircle(-0.5, 0.7, 5)
    return CF.make_system("small", "bcc", 5.42, (1, 1, 1), rcut, lmax, e, w, nspecies=2,
                          seed=seed, n_spin_cant=n_spin_cant, n_spin_pola=n_spin_pola,
                          rsteps=rsteps)


@needs_ref
@pytest.mark.parametrize("nsc,nsp,ispin,rsteps", [(1, 2, 0, None), (1, 2, 1, None),
                                                   (2, 2, 0, None), (1, 1, 0, [4.8, 5.5, 9, 9]),
                                                   (2, 2, 0, [4.8, 9, 9, 9])])
def test_build_kkr_spin_and_lmax_variants(nsc, nsp, ispin, rsteps):
    s = small_system(nsc, nsp, rsteps)
    at = s.atoms[0]
    M = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore, s.blkSizeTmatStore,
                    2, s.prel(2), nsc, nsp, ispin)
    Mr = _ref_assemble(s, 0, 2, ispin)
    assert M.shape == Mr.shape
    assert np.abs(M - Mr).max() <= 1e-14 * np.abs(Mr).max()


# ------------------------------------------------------------------- LU / tau00
def test_zgetrf_matches_lapack():
    rng = np.random.default_rng(1)
    for n in (1, 7, 64, 130):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        lu, piv, info = O.zgetrf(a)
        lu2, piv2, info2 = sl.lapack.zgetrf(a)
        assert info == info2 == 0
        assert np.array_equal(piv - 1, piv2)
        assert np.abs(lu - lu2).max() < 1e-12 * np.abs(lu2).max()


def test_zgetrf_singular_info():
    a = np.eye(5, dtype=complex)
    a[:, 2] = 0.0
    _, _, info = O.zgetrf(a)
    assert info == 3


def _c1_problem(a=0, ie=3):
    s = CF.config_c1()
    at = s.atoms[a]
    M = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore, s.blkSizeTmatStore,
                    ie, s.prel(ie))
    t0 = O.t0(3, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, ie)
    return s, M, t0


def test_tau00_zgetrf_vs_lapack_solve():
    s, M, t0 = _c1_problem()
    tau, info = O.solve_tau00_zgetrf(M, t0)
    rhs = np.zeros((M.shape[0], 16), complex)
    rhs[:16] = t0
    ref = np.linalg.solve(M, rhs)[:16]
    assert info == 0
    assert np.linalg.norm(tau - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("nblk", [2, 3, 4])
def test_tau00_block_inverse_equals_zgetrf(nblk):
    """Schur identity: (M^-1)_00 = (A - B D^-1 C)^-1, the two reference routes agree."""
    s, M, t0 = _c1_problem(1, 20)
    t1, _ = O.solve_tau00_zgetrf(M, t0)
    t2, info = O.solve_tau00_zblocklu(M, t0, nblk)
    assert info == 0
    assert np.linalg.norm(t1 - t2) <= 1e-12 * np.linalg.norm(t1)


def test_tau00_single_site_limit():
    """numLIZ = 1: M = I, tau00 = t0 (calculateTauMatrix.cpp:722-723)."""
    s = small_system(rcut=1.0)
    at = s.atoms[0]
    assert at.numLIZ == 1
    M = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore, s.blkSizeTmatStore,
                    1, s.prel(1))
    assert np.array_equal(M, np.eye(16))
    t0 = O.t0(3, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, 1)
    tau, _ = O.solve_tau00_zgetrf(M, t0)
    assert np.abs(tau - t0).max() == 0.0


def test_tau00_identity_problem():
    """inversionTest type 2 (Test/DevTests/inversionTest/inversionTest.cpp:395-401):
    G = 0, t = I -> tau00 = I."""
    M = np.eye(240, dtype=complex)
    tau, _ = O.solve_tau00_zgetrf(M, np.eye(16, dtype=complex))
    assert np.abs(tau - np.eye(16)).max() == 0.0


def test_tau00_permutation_invariance():
    """tau00 does not depend on the order of the non-central LIZ sites."""
    s = CF.config_c1()
    at = s.atoms[0]
    rng = np.random.default_rng(5)
    perm = np.concatenate([[0], 1 + rng.permutation(at.numLIZ - 1)])
    ie = 7
    t0 = O.t0(3, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, ie)
    taus = []
    for pp in (np.arange(at.numLIZ), perm):
        M = O.build_kkr(3, at.LIZPos[:, pp], at.LIZlmax[pp], at.LIZStoreIdx[pp], s.tmatStore,
                        s.blkSizeTmatStore, ie, s.prel(ie))
        taus.append(O.solve_tau00_zgetrf(M, t0)[0])
    assert np.linalg.norm(taus[0] - taus[1]) <= 1e-13 * np.linalg.norm(taus[0])


def test_tau00_small_t_limit():
    """t -> 0: tau00 = t0 + O(t^2)."""
    s, M, t0 = _c1_problem()
    eps = 1e-6
    Ms = np.eye(M.shape[0]) + eps * (M - np.eye(M.shape[0]))
    tau, _ = O.solve_tau00_zgetrf(Ms, eps * t0)
    assert np.linalg.norm(tau - eps * t0) <= 1e-9 * np.linalg.norm(eps * t0)


# ------------------------------------------------------------- post-processing
def test_spin_trafo_z_axis_is_identity():
    u, ud = O.spin_trafo(np.array([0.0, 0.0, 1.0]))
    assert np.allclose(u, [1, 0, 0, 1]) and np.allclose(ud, [1, 0, 0, 1])


def test_spin_trafo_unitary():
    rng = np.random.default_rng(2)
    for _ in range(10):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        u, ud = O.spin_trafo(v)
        U = np.array([[u[0], u[2]], [u[1], u[3]]])
        UD = np.array([[ud[0], ud[2]], [ud[1], ud[3]]])
        assert np.allclose(UD, U.conj().T) and np.allclose(U @ UD, np.eye(2))


@needs_ref
def test_spin_trafo_equals_reference():
    """or_spin_trafo against Misc/spin_trafo.f compiled from source, bit for
    bit: random directions, the axes, and directions within the tolerance of
    the south pole (the branch u(1) < 1e-8)."""
    rng = np.random.default_rng(11)
    dirs = [rng.normal(size=3) for _ in range(200)]
    dirs += [np.array(v, float) for v in ((0, 0, 1), (0, 0, -1), (1, 0, 0), (0, 1, 0),
                                          (0, -1, 0), (-1, 0, 0))]
    dirs += [np.array([1e-9, 2e-9, -1.0]), np.array([3e-7, -1e-7, -1.0])]
    for v in dirs:
        v = v / np.linalg.norm(v)
        u, ud = O.spin_trafo(v)
        ur, udr = R.spin_trafo(v)
        assert np.array_equal(u, ur) and np.array_equal(ud, udr), v


def test_trgtol_brute_force():
    """ml = (U (x) 1) mg (U^dagger (x) 1), stored as the 4 spin blocks."""
    rng = np.random.default_rng(4)
    nr = nc = 16
    mg = rng.normal(size=(2 * nr, 2 * nc)) + 1j * rng.normal(size=(2 * nr, 2 * nc))
    v = rng.normal(size=3)
    u, ud = O.spin_trafo(v / np.linalg.norm(v))
    ml = O.trgtol(nr, nc, u, ud, mg)
    U = np.array([[u[0], u[2]], [u[1], u[3]]])
    UD = np.array([[ud[0], ud[2]], [ud[1], ud[3]]])
    full = np.kron(U, np.eye(nr)) @ mg @ np.kron(UD, np.eye(nc))
    for k, (s, sp_) in enumerate([(0, 0), (1, 0), (0, 1), (1, 1)]):
        assert np.abs(ml[:, :, k] - full[s * nr:(s + 1) * nr, sp_ * nc:(sp_ + 1) * nc]).max() < 1e-13


def test_local_global_t_roundtrip():
    """The configs' global-frame t rotates back to diag(tu, td) under trgtol."""
    rng = np.random.default_rng(6)
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    tu = np.diag(rng.normal(size=16) + 1j * rng.normal(size=16))
    td = np.diag(rng.normal(size=16) + 1j * rng.normal(size=16))
    tg = CF.local_to_global_t(tu, td, v)
    u, ud = O.spin_trafo(v)
    ml = O.trgtol(16, 16, u, ud, tg)
    assert np.abs(ml[:, :, 0] - tu).max() < 1e-13 and np.abs(ml[:, :, 3] - td).max() < 1e-13
    assert np.abs(ml[:, :, 1]).max() < 1e-13 and np.abs(ml[:, :, 2]).max() < 1e-13


# ---------------------------------------------------------------------- contour
def test_congauss_integrates_analytic_functions():
    """Arc quadrature equals the straight-line integral for analytic f."""
    e, w = O.congauss(-0.5, 0.7, 31)
    for f, F in ((lambda z: z ** 2, lambda z: z ** 3 / 3), (np.exp, np.exp)):
        assert abs(np.sum(w * f(e)) - (F(0.7) - F(-0.5))) < 1e-13
    assert abs(e[0].real + 0.5) < 0.01 and abs(e[-1].real - 0.7) < 0.01
    assert np.all(e.imag > 0)


@needs_ref
def test_congauss_vs_reference():
    e, w = O.congauss(-0.5, 0.7, 31)
    e2, w2 = R.congauss(-0.5, 0.7, 0.0, 31)
    assert len(e2) == 32 and w2[31] == 0
    assert np.array_equal(e, e2[:31]) and np.array_equal(w, w2[:31])


# ------------------------------------------------- relativistic KKR blocks
@needs_ref
def test_clebsch_equals_compiled_reference():
    """Misc/clebsch.f: the oracle's tables equal the common block /cgc/ the
    compiled routine fills, bit for bit, all 500 entries."""
    a = O.clebsch()
    b = R.clebsch()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@needs_ref
@pytest.mark.parametrize("lmax", [0, 1, 2, 3])
def test_relmtrx_equals_compiled_reference(lmax):
    rng = np.random.default_rng(lmax + 40)
    kkr = (lmax + 1) ** 2
    a = rng.normal(size=(kkr, kkr)) + 1j * rng.normal(size=(kkr, kkr))
    assert np.array_equal(O.relmtrx(a), R.relmtrx(a))


@pytest.mark.parametrize("lmax", [0, 1, 2, 3])
def test_clebsch_matrix_orthogonal_and_relmtrx_similarity(lmax):
    """U is real orthogonal (a change of basis (l m s) -> (kappa mu)), each
    kappa-mu row has the norm of a Clebsch-Gordan column, and relmtrx(a) is
    U diag_spin(a) U^T."""
    U = O.clebsch_matrix(lmax)
    kkr = (lmax + 1) ** 2
    assert np.abs(U.T @ U - np.eye(2 * kkr)).max() < 1e-15
    rng = np.random.default_rng(lmax)
    a = rng.normal(size=(kkr, kkr)) + 1j * rng.normal(size=(kkr, kkr))
    ds = np.kron(np.eye(2), a)
    assert np.abs(O.relmtrx(a) - U @ ds @ U.T).max() < 1e-14 * np.abs(a).max()


def test_clebsch_closed_form_j_blocks():
    """Within each l the first 2l kappa-mu rows (j = l - 1/2) and the next
    2l + 2 (j = l + 1/2) span the l block of (m, s) states: the diagonal
    blocks of U diag(l) U^T are l on every row, and a scalar (l-independent)
    operator stays scalar."""
    lmax = 3
    U = O.clebsch_matrix(lmax)
    kkr = (lmax + 1) ** 2
    lvec = np.array([math.isqrt(k) for k in range(kkr)])
    L = U @ np.diag(np.concatenate([lvec, lvec])) @ U.T
    rows = np.concatenate([[l] * (4 * l + 2) for l in range(lmax + 1)])
    assert np.abs(L - np.diag(rows)).max() < 1e-14


@needs_ref
def test_build_kkr_rel_vs_reference_blocks():
    """The relativistic M, block by block from the compiled makegij_ and the
    relativistic branch of setgij_ (relmtrx, prel^2/E) with the full kappa-mu
    t of each site (calculateTauMatrix.cpp:216-242)."""
    s = CF.config_rel()
    at = s.atoms[0]
    ie = 2
    e, p = s.energies[ie], s.prel(ie)
    M = O.build_kkr_rel(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore,
                        s.blkSizeTmatStore, ie, p, e)
    n = at.numLIZ
    Mr = np.eye(32 * n, dtype=complex)
    cg = R.cgaunt(3)
    il, i1 = R.ifacts(3)
    blk = s.blkSizeTmatStore
    for i in range(n):
        t = s.tmatStore[ie * blk:(ie + 1) * blk, at.LIZStoreIdx[i]].reshape(32, 32, order="F")
        for j in range(n):
            if i == j:
                continue
            g = R.makegij(3, 3, 3, p, at.LIZPos[:, i] - at.LIZPos[:, j], cg, il, i1)
            Mr[32 * i:32 * i + 32, 32 * j:32 * j + 32] = -t @ R.setgij_rel(g, e, p)
    assert np.abs(M - Mr).max() <= 1e-14 * np.abs(Mr).max()


def test_relativistic_matrix_is_a_similarity_of_the_spin_canted_one():
    """M_rel = U~ M' U~^T with M' the spin-canted matrix of t' = (prel^2/E)
    U^T t U, and tau00_rel = (E/prel^2) U tau00'(t0') U^T against LAPACK on
    M_rel: the identity the device path computes with."""
    s = CF.config_rel()
    at = s.atoms[1]
    ie = 1
    e, p = s.energies[ie], s.prel(ie)
    f = p * p / e
    U = O.clebsch_matrix(3)
    Mrel = O.build_kkr_rel(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore,
                           s.blkSizeTmatStore, ie, p, e)
    blk = s.blkSizeTmatStore
    tp = np.array(s.tmatStore, order="F", copy=True)
    for col in range(tp.shape[1]):
        for k in range(len(s.energies)):
            t = tp[k * blk:(k + 1) * blk, col].reshape(32, 32, order="F")
            fk = s.prel(k) ** 2 / s.energies[k]
            tp[k * blk:(k + 1) * blk, col] = (fk * U.T @ t @ U).flatten(order="F")
    Mp = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, tp, blk, ie, p, 2, 2)
    Ut = np.kron(np.eye(at.numLIZ), U)
    assert np.abs(Ut @ Mp @ Ut.T - Mrel).max() < 1e-13 * np.abs(Mrel).max()
    t0 = s.tmatStore[ie * blk:(ie + 1) * blk, at.LIZStoreIdx[0]].reshape(32, 32, order="F")
    ref, _ = O.solve_tau00_zgetrf(Mrel, t0)
    t0p = tp[ie * blk:(ie + 1) * blk, at.LIZStoreIdx[0]].reshape(32, 32, order="F")
    taup, _ = O.solve_tau00_zgetrf(Mp, t0p)
    got = U @ taup @ U.T / f
    assert np.linalg.norm(got - ref) < 1e-12 * np.linalg.norm(ref)


def test_tripmt_is_the_triple_product():
    rng = np.random.default_rng(8)
    u, b, v = (rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32)) for _ in range(3))
    assert np.abs(O.tripmt(u, b, v) - u @ b @ v).max() < 1e-12 * np.abs(u @ b @ v).max()
