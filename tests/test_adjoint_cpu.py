"""CPU checks of the inputs and references of the adjoint route (no GPU).

They pin the pieces the GPU tests lean on: dt/dE of the synthetic store,
d kappa / dE, the two independent routes to d ln det M / dE on the oracle's
matrices, and the C5 geometry.
"""
import numpy as np

from lsms_amd import configs as CF
from oracle import oracle as O
from tests import adjoint_ref as R


def small(nsc=1, nsp=1, rsteps=None, ne=4):
    e, w = CF.contour_semicircle(-0.5, 0.7, ne)
    return CF.make_system("small", "bcc", 5.42, (1, 1, 1), 6.0, 3, e, w, nspecies=2, seed=5,
                          n_spin_cant=nsc, n_spin_pola=nsp, rsteps=rsteps)


def test_dprel_matches_difference():
    for e in [0.3 + 0.2j, -0.4 + 0.05j, 0.69 + 0.001j]:
        h = 1e-6
        fd = (CF.prel_of(e + h) - CF.prel_of(e - h)) / (2 * h)
        assert abs(CF.dprel_of(e) - fd) < 1e-8 * abs(fd)


def test_dtmat_store_matches_difference():
    for nsc, nsp in [(1, 1), (1, 2), (2, 2)]:
        s = small(nsc, nsp)
        d = s.dtmat_store()
        h = 1e-6
        for ie, e in enumerate(s.energies[:2]):
            sp = s.at_energies([e + h, e - h])
            blk = s.blkSizeTmatStore
            fd = (sp.tmatStore[:blk] - sp.tmatStore[blk:2 * blk]) / (2 * h)
            ref = d[ie * blk:(ie + 1) * blk]
            assert np.max(np.abs(fd - ref)) < 1e-7 * np.max(np.abs(ref))


def test_at_energies_reproduces_store():
    s = small(2, 2)
    s2 = s.at_energies(s.energies)
    assert np.array_equal(s2.tmatStore, s.tmatStore)


def test_lloyd_two_routes_agree_on_oracle():
    """Tr(M^-1 dM/dE) (four-point derivative) = d ln det M / dE (difference of
    LAPACK log-determinants): the identity the GPU Lloyd term is checked with."""
    for nsc, nsp, ispin in [(1, 1, 0), (1, 2, 1), (2, 2, 0)]:
        s = small(nsc, nsp)
        for e in s.energies[:2]:
            tr = R.lloyd_trace(s, 0, e, ispin)
            fd = R.lloyd_logdet_fd(s, 0, e, ispin)
            assert abs(tr - fd) < 1e-6 * abs(tr), (nsc, nsp, e, tr, fd)


def test_lloyd_trace_detects_a_dropped_term():
    """The trace check is sensitive to each part of dM/dE: dropping the dt/dE
    part (t held at E) changes it by far more than the tolerance."""
    s = small()
    e = s.energies[1]
    tr = R.lloyd_trace(s, 0, e)
    # dM/dE with t frozen: -t dG/dE only
    h = 1e-3
    s_e = s.at_energies([e])
    m = R.oracle_m(s_e, 0, 0)

    def m_frozen(ep):  # t at E, G at ep
        at = s.atoms[0]
        return O.build_kkr(s.maxlmax, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s_e.tmatStore,
                           s.blkSizeTmatStore, 0, complex(CF.prel_of(ep)))

    d = (m_frozen(e + h) - m_frozen(e - h) - 1j * (m_frozen(e + 1j * h) - m_frozen(e - 1j * h))) / (4 * h)
    tr_frozen = complex(np.trace(np.linalg.solve(m, d)))
    assert abs(tr_frozen - tr) > 1e-3 * abs(tr)


def test_c5_geometry():
    s = CF.config_c5(npts=2)
    assert s.n_spin_cant == 2 and s.n_spin_pola == 2
    assert all(a.numLIZ == 137 for a in s.atoms)
    assert all(s.nrmat_ns(i) == 4384 for i in range(len(s.atoms)))
    assert len(s.atoms) == 8
    assert s.tmatStore.shape[1] == 16  # one t-matrix type per site of the 16-site cell
    assert np.allclose(np.linalg.norm(s.evecs, axis=1), 1.0)
    # LIZ store index is the site itself, central first
    assert all(int(a.LIZStoreIdx[0]) == i for i, a in enumerate(s.atoms))


def test_idos_formulas_on_closed_forms():
    """The contour weights integrate analytic functions along the semicircle:
    a constant tau00 - t = c I (k x k) gives -(1/pi) Im(c k (etop - ebot)),
    and a Lloyd term f(E) = 1/(E - z0) with z0 below the axis gives
    -(1/pi) Im ln((etop - z0)/(ebot - z0))."""
    from lsms_amd import contour
    e, w = CF.contour_semicircle(-0.5, 0.7, 31)
    c = 0.3 - 0.2j
    k = 16
    tau = np.array([c * np.eye(k) + np.eye(k) for _ in e])
    t0 = np.array([np.eye(k) for _ in e])
    assert abs(contour.idos_ms(tau, t0, w) - (-np.imag(c * k * 1.2) / np.pi)) < 1e-12
    z0 = 0.1 - 0.3j
    f = 1.0 / (e - z0)
    ref = -np.imag(np.log((0.7 - z0) / (-0.5 - z0))) / np.pi
    assert abs(contour.idos_lloyd(f, w) - ref) < 1e-10
