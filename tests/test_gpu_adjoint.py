"""GPU tests of the explicit-inverse routes through the C ABI: full tau
(solveTauFullzgetrf), the Lloyd term Tr(M^-1 dM/dE) and the tau00 gradient
with respect to the t-matrix store, plus the torch.autograd wrapper.

Tolerances:
  - full tau and tau00: 1e-10 relative (Frobenius), the FP64 bar;
  - Lloyd term against the oracle's Tr(M^-1 dM/dE) (four-point derivative,
    error O(h^4) = 1e-12): 1e-8 relative; against the difference of LAPACK
    log-determinants (h = 1e-6 Ry): 1e-6 relative;
  - gradient against central differences of the FP64 tau00 along random
    complex store directions (h = 1e-6): 1e-6 relative.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import torch

from lsms_amd import capi
from lsms_amd import configs as CF
from lsms_amd import torch_adjoint as TA
from tests import adjoint_ref as R

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def small(nsc=1, nsp=1, rsteps=None, ne=4, seed=5, rcut=6.0):
    e, w = CF.contour_semicircle(-0.5, 0.7, ne)
    return CF.make_system("small", "bcc", 5.42, (1, 1, 1), rcut, 3, e, w, nspecies=2, seed=seed,
                          n_spin_cant=nsc, n_spin_pola=nsp, rsteps=rsteps)


def engine(s):
    ctx = capi.Context(0)
    ctx.load_system(s)
    ctx.upload_dtmats(s.dtmat_store(), s.blkSizeTmatStore, len(s.energies))
    return ctx


@pytest.fixture(scope="module")
def c1():
    s = CF.config_c1()
    return s, engine(s)


def test_tau_full_c1(c1):
    s, ctx = c1
    for a, ie in [(0, 0), (1, 15), (0, 30)]:
        tau, info = ctx.tau_full(a, ie, s.prel(ie), s.nrmat_ns(a))
        assert info == 0
        assert rel(tau, R.tau_full(s, a, ie)) < 1e-10


@pytest.mark.parametrize("nsc,nsp,ispin,rsteps", [(2, 2, 0, None), (1, 2, 1, None),
                                                  (2, 2, 0, [5.0, 5.5, 100.0, 100.0])])
def test_tau_full_variants(nsc, nsp, ispin, rsteps):
    s = small(nsc, nsp, rsteps)
    ctx = engine(s)
    ie = 2
    tau, info = ctx.tau_full(0, ie, s.prel(ie), s.nrmat_ns(0), ispin=ispin)
    assert info == 0
    assert rel(tau, R.tau_full(s, 0, ie, ispin)) < 1e-10


def test_adjoint_tau00_equals_batched_c1(c1):
    s, ctx = c1
    n = len(s.atoms) * len(s.energies)
    atoms = np.repeat(np.arange(len(s.atoms)), len(s.energies))
    iies = np.tile(np.arange(len(s.energies)), len(s.atoms))
    tb, _ = ctx.tau00_batched(atoms, iies, [s.prel(i) for i in iies], s.kkrsz_ns)
    ta, ll, _, info = ctx.tau00_adjoint(atoms, iies, s.energies[iies], s.kkrsz_ns)
    assert np.all(info == 0)
    assert max(rel(ta[b], tb[b]) for b in range(n)) < 1e-10
    assert np.all(np.isfinite(ll))


def test_lloyd_c1_against_oracle(c1):
    s, ctx = c1
    cases = [(0, 0), (0, 12), (1, 20), (1, 30)]
    atoms = [a for a, _ in cases]
    iies = [ie for _, ie in cases]
    _, ll, _, info = ctx.tau00_adjoint(atoms, iies, s.energies[iies], s.kkrsz_ns)
    assert np.all(info == 0)
    for b, (a, ie) in enumerate(cases):
        e = s.energies[ie]
        tr = R.lloyd_trace(s, a, e)
        fd = R.lloyd_logdet_fd(s, a, e)
        assert abs(ll[b] - tr) < 1e-8 * abs(tr), (a, ie, ll[b], tr)
        assert abs(ll[b] - fd) < 1e-6 * abs(tr), (a, ie, ll[b], fd)


@pytest.mark.parametrize("nsc,nsp,ispin,rsteps", [(1, 2, 0, None), (1, 2, 1, None),
                                                  (2, 2, 0, None),
                                                  (2, 2, 0, [5.0, 5.5, 100.0, 100.0]),
                                                  (1, 1, 0, [5.0, 100.0, 100.0, 100.0])])
def test_lloyd_variants(nsc, nsp, ispin, rsteps):
    s = small(nsc, nsp, rsteps)
    ctx = engine(s)
    iies = [0, 3]
    _, ll, _, info = ctx.tau00_adjoint([0, 1], iies, s.energies[iies], s.kkrsz_ns,
                                       ispins=[ispin, ispin])
    assert np.all(info == 0)
    for b, (a, ie) in enumerate(zip([0, 1], iies)):
        tr = R.lloyd_trace(s, a, s.energies[ie], ispin)
        assert abs(ll[b] - tr) < 1e-8 * abs(tr), (a, ie, ll[b], tr)


def tau_batch(ctx, s, store, atoms, iies, ispins):
    ctx.upload_tmats(store, s.blkSizeTmatStore, len(s.energies))
    tau, info = ctx.tau00_batched(atoms, iies, [s.prel(i) for i in iies], s.kkrsz_ns,
                                  ispins=ispins)
    assert np.all(info == 0)
    return tau


@pytest.mark.parametrize("nsc,nsp,ispin,rsteps", [(1, 1, 0, None), (1, 2, 1, None),
                                                  (2, 2, 0, None),
                                                  (2, 2, 0, [5.0, 5.5, 100.0, 100.0])])
def test_tau_gradient_against_differences(nsc, nsp, ispin, rsteps):
    s = small(nsc, nsp, rsteps)
    ctx = engine(s)
    rng = np.random.default_rng(11)
    atoms, iies = [0, 1, 0], [0, 2, 3]
    ispins = [ispin] * 3
    k = s.kkrsz_ns
    V = rng.normal(size=(3, k, k)) + 1j * rng.normal(size=(3, k, k))
    _, _, g, info = ctx.tau00_adjoint(atoms, iies, s.energies[iies], k, ispins=ispins, lloyd=False,
                                      tau00_bar=V, store_shape=s.tmatStore.shape)
    assert np.all(info == 0)
    h = 1e-6
    for trial in range(2):
        D = rng.normal(size=s.tmatStore.shape) + 1j * rng.normal(size=s.tmatStore.shape)
        D *= np.abs(s.tmatStore) > 0  # perturb the entries the store holds
        for direction in (D, 1j * D):
            tp = tau_batch(ctx, s, s.tmatStore + h * direction, atoms, iies, ispins)
            tm = tau_batch(ctx, s, s.tmatStore - h * direction, atoms, iies, ispins)
            fd = np.real(np.vdot(V, (tp - tm) / (2 * h)))
            an = np.real(np.vdot(g, direction))
            assert abs(an - fd) < 1e-6 * (abs(fd) + 1e-3 * np.linalg.norm(g)), (trial, an, fd)
    ctx.upload_tmats(s.tmatStore, s.blkSizeTmatStore, len(s.energies))


def test_torch_autograd_wrapper():
    s = small(2, 2)
    ctx = engine(s)
    atoms, iies = [0, 1], [1, 2]
    store = torch.tensor(s.tmatStore, dtype=torch.complex128, requires_grad=True)
    rng = np.random.default_rng(3)
    W = torch.tensor(rng.normal(size=(2, 32, 32)) + 1j * rng.normal(size=(2, 32, 32)))
    tau = TA.tau00(store, ctx, s, atoms, iies)
    loss = torch.real(torch.sum(torch.conj(W) * tau))
    loss.backward()
    g = store.grad.numpy()
    # same value as the plain batched path
    tb = tau_batch(ctx, s, s.tmatStore, atoms, iies, None)
    assert rel(tau.detach().numpy(), tb) < 1e-10
    # directional derivative of the real loss = Re <grad, direction>
    h = 1e-6
    D = rng.normal(size=s.tmatStore.shape) * (np.abs(s.tmatStore) > 0)
    lp = np.real(np.sum(np.conj(W.numpy()) * tau_batch(ctx, s, s.tmatStore + h * D, atoms, iies, None)))
    lm = np.real(np.sum(np.conj(W.numpy()) * tau_batch(ctx, s, s.tmatStore - h * D, atoms, iies, None)))
    fd = (lp - lm) / (2 * h)
    assert abs(np.real(np.vdot(g, D)) - fd) < 1e-6 * abs(fd)


def test_c5_full_size_adjoint():
    """Spin-canted bcc 137-site LIZ (N = 4384): tau00 against LAPACK and the
    Lloyd term against the oracle's trace, for two (atom, energy) solves."""
    s = CF.config_c5()
    ctx = engine(s)
    cases = [(0, 3), (5, 27)]
    atoms = [a for a, _ in cases]
    iies = [ie for _, ie in cases]
    ta, ll, _, info = ctx.tau00_adjoint(atoms, iies, s.energies[iies], 32)
    assert np.all(info == 0)
    tb, _ = ctx.tau00_batched(atoms, iies, [s.prel(i) for i in iies], 32)
    for b, (a, ie) in enumerate(cases):
        m = R.oracle_m(s, a, ie)
        assert m.shape == (4384, 4384)
        t0 = R.oracle_t0(s, a, ie)
        rhs = np.zeros((4384, 32), dtype=np.complex128)
        rhs[:32] = t0
        ref = sla.solve(m, rhs)[:32]
        assert rel(ta[b], ref) < 1e-10
        assert rel(tb[b], ref) < 1e-10
        tr = R.lloyd_trace(s, a, s.energies[ie])
        assert abs(ll[b] - tr) < 1e-8 * abs(tr), (a, ie, ll[b], tr)


def test_c5_full_size_gradient_and_lloyd_by_differences():
    """N = 4384 (spin-canted bcc 137 sites), one solve: the tau00 gradient
    against central differences of the GPU's FP64 tau00 along a random
    complex store direction and its i-rotation (h = 1e-5: truncation
    O(h^2) ~ 1e-10, rounding ~1e-12 / h ~ 1e-7 relative), and the Lloyd
    term against the central difference of LAPACK log-determinants of the
    oracle's M(E +- h), h = 1e-6 Ry."""
    s = CF.config_c5(natoms=1)
    ctx = engine(s)
    a, ie = 0, 11
    k = 32
    rng = np.random.default_rng(29)
    V = rng.normal(size=(1, k, k)) + 1j * rng.normal(size=(1, k, k))
    _, ll, g, info = ctx.tau00_adjoint([a], [ie], s.energies[[ie]], k, tau00_bar=V,
                                       store_shape=s.tmatStore.shape)
    assert np.all(info == 0)
    h = 1e-5
    D = rng.normal(size=s.tmatStore.shape) + 1j * rng.normal(size=s.tmatStore.shape)
    D *= np.abs(s.tmatStore) > 0
    for direction in (D, 1j * D):
        tp = tau_batch(ctx, s, s.tmatStore + h * direction, [a], [ie], None)
        tm = tau_batch(ctx, s, s.tmatStore - h * direction, [a], [ie], None)
        fd = np.real(np.vdot(V, (tp - tm) / (2 * h)))
        an = np.real(np.vdot(g, direction))
        print(f"C5 gradient: adjoint {an:.10e} differences {fd:.10e}")
        assert abs(an - fd) < 1e-6 * abs(fd), (an, fd)
    ctx.upload_tmats(s.tmatStore, s.blkSizeTmatStore, len(s.energies))
    ctx.close()
    fd_ll = R.lloyd_logdet_fd(s, a, s.energies[ie], h=1e-6)
    print(f"C5 Lloyd: adjoint {ll[0]} log-det differences {fd_ll}")
    assert abs(ll[0] - fd_ll) < 1e-6 * abs(fd_ll), (ll[0], fd_ll)
