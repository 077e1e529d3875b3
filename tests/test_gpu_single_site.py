"""Single-site scattering on the GPU (lsmsx_single_site_nonrel) against the
oracle's restatement of single_scatterer_nonrel.f / semrel_c.f -- needs a GPU.

The kernels follow the same fit, Bulirsch-Stoer steps and matching as the
oracle in FP64; floating-point contraction on the GPU can move an adaptive
step decision, so the bar is relative: 1e-10 on matom and on every zlr/jlr
column (relative to the column's largest value), the accuracy bar of the
tau00 path these t-matrices feed.
"""
import numpy as np
import pytest

from lsms_amd import capi
from lsms_amd import configs as CF
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def colerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("nrelv", [0, 10])
def test_batch_matches_oracle(ctx, nrelv):
    rr, vr, jm, jw, rs, e, pr = CF.single_site_batch(nspecies=4, ne=48)
    print(f"single-site batch nrelv={nrelv} npot={len(rr)} ne={len(e)}", flush=True)
    matom, tmat, zlr, jlr, rc = ctx.single_site_nonrel(nrelv, 3, vr, rr, jm, jw, rs, e, pr)
    assert rc == 0
    worst = [0.0, 0.0, 0.0]
    for q in range(len(rr)):
        n = jw[q]
        for k in range(len(e)):
            m, t, z, j, orc = O.single_scatterer_nonrel(nrelv, 3, e[k], pr[k], vr[q], rr[q], jm[q],
                                                        jw[q], rs[q])
            assert orc == 0
            worst[0] = max(worst[0], np.max(np.abs(matom[k, q] - m) / np.abs(m)))
            assert np.abs(tmat[k, q] - t).max() <= TOL * np.abs(t).max()
            for l in range(4):
                worst[1] = max(worst[1], colerr(zlr[k, q, l, :n], z[:n, l]))
                worst[2] = max(worst[2], colerr(jlr[k, q, l, :n], j[:n, l]))
            assert not np.any(zlr[k, q, :, n:]) and not np.any(jlr[k, q, :, n:])
    print("max rel err matom/zlr/jlr:", worst, flush=True)
    assert max(worst) < TOL, worst


def test_constant_potential_closed_form_on_gpu(ctx):
    """Constant potential inside the sphere, matched on a mesh point: matom
    against the spherical-Bessel closed form (the oracle's own pin)."""
    from tests.test_single_site_cpu import square_well_matom
    rmt, jmt = 2.3, 400
    r = CF.log_mesh(rmt, jmt)
    rs = r[jmt - 1]
    es = np.array([0.6 + 0.05j, -0.3 + 0.4j, 1.2 + 0.001j])
    vr = np.stack([-0.8 * r, 0.3 * r])
    matom, _, _, _, rc = ctx.single_site_nonrel(10, 3, vr, np.stack([r, r]), jmt, jmt + 15, rs,
                                                es, np.sqrt(es), radial=False)
    assert rc == 0
    for q, V0 in enumerate((-0.8, 0.3)):
        for k, E in enumerate(es):
            ref = square_well_matom(V0, E, rs, 3)
            assert np.max(np.abs(matom[k, q] - ref) / np.abs(ref)) < 1e-12


def test_repeat_and_subset_are_identical(ctx):
    """A second call, and a call on a subset of the potentials and energies,
    give bitwise the same numbers (one thread per (energy, potential, l))."""
    rr, vr, jm, jw, rs, e, pr = CF.single_site_batch(nspecies=2, ne=8)
    a = ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr)
    b = ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr)
    c = ctx.single_site_nonrel(0, 3, vr[1:3], rr[1:3], jm[1:3], jw[1:3], rs[1:3], e[2:5], pr[2:5])
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    assert np.array_equal(c[0], a[0][2:5, 1:3])
    assert np.array_equal(c[2], a[2][2:5, 1:3])


def test_argument_errors(ctx):
    r, vr, jmt, jws, rs = CF.model_potential(26, 2.3, iprpts=465)
    e = np.array([0.5 + 0.01j])
    with pytest.raises(capi.LsmsxError):  # jmt + 11 > iprpts
        ctx.single_site_nonrel(0, 3, vr, r, jmt, jws, rs, e, np.sqrt(e))
    r, vr, jmt, jws, rs = CF.model_potential(26, 2.3)
    with pytest.raises(capi.LsmsxError):  # r_sph outside the mesh
        ctx.single_site_nonrel(0, 3, vr, r, jmt, jws, -1.0, e, np.sqrt(e))
    with pytest.raises(capi.LsmsxError):  # lmax above the path's 3
        ctx.single_site_nonrel(0, 4, vr, r, jmt, jws, rs, e, np.sqrt(e))
    m, _, _, _, rc = ctx.single_site_nonrel(0, 3, vr, r, jmt, jws, rs, e[:0], e[:0])
    assert rc == 0 and m.shape == (0, 1, 4)


def test_gpu_tmatrices_feed_tau00(ctx):
    """End to end on the GPU: t-matrices of two model potentials (Fe- and
    Cu-like, scalar-relativistic) at the bcc-15 contour energies fill the
    t-matrix store in place of the synthetic phase shifts; tau00 of every
    atom at a spread of energies then meets the path's 1e-10 bar against
    LAPACK on the oracle's KKR matrices built from the same store, and the
    store itself matches the oracle's single-site t to 1e-10."""
    import scipy.linalg as sla
    from tests import adjoint_ref as R
    s = CF.config_c1()
    pots = [CF.model_potential(26, 2.3), CF.model_potential(29, 2.4, screen=0.65)]
    rr = np.stack([p[0] for p in pots])
    vr = np.stack([p[1] for p in pots])
    jm = [p[2] for p in pots]
    jw = [p[3] for p in pots]
    rs = [p[4] for p in pots]
    prels = np.array([s.prel(i) for i in range(len(s.energies))])
    matom, tmat, _, _, rc = ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, s.energies, prels,
                                                   radial=False)
    assert rc == 0
    kk = 16
    store = np.zeros_like(s.tmatStore)
    for sp in range(2):
        for ie in range(len(s.energies)):
            base = ie * s.blkSizeTmatStore
            store[base:base + kk * kk, sp] = tmat[ie, sp].flatten(order="F")
            if ie % 10 == 0:
                t_or = O.single_scatterer_nonrel(0, 3, s.energies[ie], prels[ie], vr[sp], rr[sp],
                                                 jm[sp], jw[sp], rs[sp])[1]
                assert np.abs(tmat[ie, sp] - t_or).max() <= TOL * np.abs(t_or).max()
    s.tmatStore = store
    c2 = capi.Context(0)
    c2.load_system(s)
    na = len(s.atoms)
    iies = [0, 7, 15, 23, 30]
    atoms = np.repeat(np.arange(na), len(iies))
    ies = np.tile(iies, na)
    tau, info = c2.tau00_batched(atoms, ies, [s.prel(i) for i in ies], kk)
    c2.close()
    assert np.all(info == 0)
    for b in range(len(atoms)):
        m = R.oracle_m(s, atoms[b], ies[b])
        rhs = np.zeros((m.shape[0], kk), dtype=np.complex128)
        rhs[:kk] = R.oracle_t0(s, atoms[b], ies[b])
        ref = sla.solve(m, rhs)[:kk]
        assert np.linalg.norm(tau[b] - ref) / np.linalg.norm(ref) < 1e-10, (atoms[b], ies[b])


def test_green_mt_matches_oracle(ctx):
    """Muffin-tin DOS and green(r) on the GPU against the oracle, for radial
    functions of four potentials at six energies and random tau00 near t,
    both spin conventions; the GPU quadrature uses the same ring of fitted
    points, so the bar is 1e-12 relative."""
    rr, vr, jm, jw, rs, e, pr = CF.single_site_batch(nspecies=2, ne=6)
    npot, ne = len(rr), len(e)
    matom, tmat, zlr, jlr, rc = ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr)
    assert rc == 0
    rng = np.random.default_rng(11)
    nsolve = npot * ne
    pot_of = np.tile(np.arange(npot), ne)
    z = zlr.reshape(nsolve, 4, -1)
    j = jlr.reshape(nsolve, 4, -1)
    tau = tmat.reshape(nsolve, 16, 16) + 0.05 * (rng.normal(size=(nsolve, 16, 16)) +
                                                1j * rng.normal(size=(nsolve, 16, 16)))
    for nspol in (1, 2):
        dos, green = ctx.green_mt(3, tau, z, j, pot_of, rr, jm, jw, rs, n_spin_pola=nspol)
        for b in range(nsolve):
            p = pot_of[b]
            d_or, g_or, _, _ = O.green_mt(3, tau[b], z[b].T, j[b].T, rr[p], jm[p], jw[p], rs[p],
                                          nspol)
            assert abs(dos[b] - d_or) <= 1e-12 * abs(d_or), (b, dos[b], d_or)
            n = jw[p]
            assert np.abs(green[b, :n] - g_or).max() <= 1e-12 * np.abs(g_or).max()
            assert not np.any(green[b, n:])


def test_dos_chain_on_gpu(ctx):
    """The whole chain on the GPU for the bcc-15 system with model
    potentials: single-site t and radial functions, tau00 of every atom at
    five energies, then the muffin-tin DOS; against the oracle chain
    (oracle t, LAPACK tau00 on the oracle's KKR matrices, oracle DOS) to
    1e-10 relative."""
    import scipy.linalg as sla
    from tests import adjoint_ref as R
    s = CF.config_c1()
    pots = [CF.model_potential(26, 2.3), CF.model_potential(29, 2.4, screen=0.65)]
    rr = np.stack([p[0] for p in pots])
    vr = np.stack([p[1] for p in pots])
    jm = np.array([p[2] for p in pots])
    jw = np.array([p[3] for p in pots])
    rs = np.array([p[4] for p in pots])
    prels = np.array([s.prel(i) for i in range(len(s.energies))])
    matom, tmat, zlr, jlr, rc = ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, s.energies, prels)
    assert rc == 0
    kk = 16
    store = np.zeros_like(s.tmatStore)
    for sp in range(2):
        for ie in range(len(s.energies)):
            base = ie * s.blkSizeTmatStore
            store[base:base + kk * kk, sp] = tmat[ie, sp].flatten(order="F")
    s.tmatStore = store
    c2 = capi.Context(0)
    c2.load_system(s)
    na = len(s.atoms)
    iies = [0, 15, 30]
    atoms = np.repeat(np.arange(na), len(iies))
    ies = np.tile(iies, na)
    tau, info = c2.tau00_batched(atoms, ies, [s.prel(i) for i in ies], kk)
    assert np.all(info == 0)
    sp_of = np.array([int(s.atoms[a].LIZStoreIdx[0]) for a in atoms])
    z = np.stack([zlr[ies[b], sp_of[b]] for b in range(len(atoms))])
    j = np.stack([jlr[ies[b], sp_of[b]] for b in range(len(atoms))])
    dos, _ = c2.green_mt(3, tau, z, j, sp_of, rr, jm, jw, rs, with_green=False)
    c2.close()
    for b in range(len(atoms)):
        a, ie, sp = atoms[b], ies[b], sp_of[b]
        m = R.oracle_m(s, a, ie)
        rhs = np.zeros((m.shape[0], kk), dtype=np.complex128)
        rhs[:kk] = R.oracle_t0(s, a, ie)
        tref = sla.solve(m, rhs)[:kk]
        _, _, zo, jo, _ = O.single_scatterer_nonrel(0, 3, s.energies[ie], prels[ie], vr[sp], rr[sp],
                                                    jm[sp], jw[sp], rs[sp])
        d_or = O.green_mt(3, tref, zo, jo, rr[sp], jm[sp], jw[sp], rs[sp], 1)[0]
        assert abs(dos[b] - d_or) <= 1e-10 * abs(d_or), (a, ie, dos[b], d_or)


def test_green_mt_resident_equals_host_path():
    """Radial functions left on the device by the single-site call give
    bitwise the Green function of the host-array call, for solves picked
    out of order; an intervening host-array call invalidates them."""
    c = capi.Context(0)
    rr, vr, jm, jw, rs, e, pr = CF.single_site_batch(nspecies=2, ne=5)
    npot = len(rr)
    with pytest.raises(capi.LsmsxError):  # nothing resident yet
        c.green_mt_resident(np.zeros((1, 16, 16)), [0], rr.shape[1])
    c.set_single_site_resident(True)
    matom, tmat, zlr, jlr, rc = c.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr)
    _, _, _, _, rc2 = c.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr, radial=False)
    assert rc == 0 and rc2 == 0
    pairs = np.array([7, 0, 19, 3, 3, 12, 15, 1, 8])
    q_e, q_p = pairs // npot, pairs % npot
    rng = np.random.default_rng(5)
    tau = tmat[q_e, q_p] + 0.05 * (rng.normal(size=(len(pairs), 16, 16)) +
                                   1j * rng.normal(size=(len(pairs), 16, 16)))
    # buffers sized for another iprpts or kkrsz are refused, not over-run
    with pytest.raises(capi.LsmsxError):
        c.green_mt_resident(tau, pairs, rr.shape[1] - 1, n_spin_pola=2)
    with pytest.raises(capi.LsmsxError):  # lmax 2 blocks against the resident lmax 3
        c.green_mt_resident(tau[:, :9, :9], pairs, rr.shape[1], n_spin_pola=2)
    with pytest.raises(ValueError):  # not a (lmax+1)^2 block, or one block short
        c.green_mt_resident(tau[:, :10, :10], pairs, rr.shape[1], n_spin_pola=2)
    with pytest.raises(ValueError):
        c.green_mt_resident(tau[1:], pairs, rr.shape[1], n_spin_pola=2)
    d_res, g_res = c.green_mt_resident(tau, pairs, rr.shape[1], n_spin_pola=2)
    # ASA: the sphere integrals run to jws (energyContourIntegration.cpp:519)
    d_res2, _ = c.green_mt_resident(tau, pairs, rr.shape[1], n_spin_pola=2, jmt_of_pot=jw)
    d_or = [O.green_mt(3, tau[b], zlr[q_e[b], q_p[b]].T, jlr[q_e[b], q_p[b]].T, rr[q_p[b]],
                       jw[q_p[b]], jw[q_p[b]], rs[q_p[b]], 2)[0] for b in range(len(pairs))]
    assert np.allclose(d_res2, d_or, rtol=1e-12, atol=0)
    assert not np.allclose(d_res2, d_res, rtol=1e-6, atol=0)
    d_host, g_host = c.green_mt(3, tau, zlr[q_e, q_p], jlr[q_e, q_p], q_p, rr, jm, jw, rs,
                                n_spin_pola=2)
    assert np.array_equal(d_res, d_host) and np.array_equal(g_res, g_host)
    with pytest.raises(capi.LsmsxError):  # the host-array call reused the buffers
        c.green_mt_resident(tau, pairs, rr.shape[1])
    c.close()


@pytest.mark.parametrize("nspol", [1, 2])
def test_store_filled_on_device(nspol):
    """The device t-matrix store filled straight from the single-site batch
    gives the tau00 of the same t uploaded from the host (to the rounding of
    1/matom on either side), and meets LAPACK on the oracle's matrices;
    collinear non-polarized and spin-polarized."""
    import scipy.linalg as sla
    from tests import adjoint_ref as R
    e, w = CF.contour_semicircle(-0.5, 0.7, 8)
    s = CF.make_system("small", "bcc", 5.42, (1, 1, 1), 6.0, 3, e, w, nspecies=2, seed=5,
                       n_spin_pola=nspol)
    pots, store_of, spin_of = [], [], []
    for sp in range(2):
        for isp in range(nspol):
            pots.append(CF.model_potential(26 + 3 * sp, 2.3, vmtz=-1.2 + 0.1 * isp))
            store_of.append(sp)
            spin_of.append(isp)
    rr = np.stack([p[0] for p in pots])
    vr = np.stack([p[1] for p in pots])
    jm = np.array([p[2] for p in pots])
    jw = np.array([p[3] for p in pots])
    rs = np.array([p[4] for p in pots])
    prels = np.array([s.prel(i) for i in range(len(s.energies))])
    c = capi.Context(0)
    c.load_system(s)
    _, tmat, _, _, rc = c.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, s.energies, prels,
                                             radial=False)
    assert rc == 0
    c.tmats_from_single_site(s.tmatStore.shape[0], s.tmatStore.shape[1], s.blkSizeTmatStore,
                             len(s.energies), store_of, spin_of)
    store = np.zeros_like(s.tmatStore)
    for p in range(len(pots)):
        for ie in range(len(s.energies)):
            base = ie * s.blkSizeTmatStore + spin_of[p] * 256
            store[base:base + 256, store_of[p]] = tmat[ie, p].flatten(order="F")
    s.tmatStore = store
    na = len(s.atoms)
    atoms = np.repeat(np.arange(na), 3 * nspol)
    ies = np.tile(np.repeat([0, 3, 7], nspol), na)
    isps = np.tile(np.arange(nspol), 3 * na)
    pr = [s.prel(i) for i in ies]
    tau_dev, info = c.tau00_batched(atoms, ies, pr, 16, ispins=isps)
    c.close()
    c2 = capi.Context(0)
    c2.load_system(s)
    tau_host, info2 = c2.tau00_batched(atoms, ies, pr, 16, ispins=isps)
    c2.close()
    assert np.all(info == 0) and np.all(info2 == 0)
    for b in range(len(atoms)):
        assert np.linalg.norm(tau_dev[b] - tau_host[b]) <= 1e-13 * np.linalg.norm(tau_host[b])
    for b in (0, len(atoms) - 1):
        m = R.oracle_m(s, atoms[b], ies[b], isps[b])
        rhs = np.zeros((m.shape[0], 16), dtype=np.complex128)
        rhs[:16] = R.oracle_t0(s, atoms[b], ies[b], isps[b])
        ref = sla.solve(m, rhs)[:16]
        assert np.linalg.norm(tau_dev[b] - ref) / np.linalg.norm(ref) < 1e-10


def test_green_mt_canted_matches_oracle(ctx):
    """Spin-canted muffin-tin DOS and green(r) on the GPU against the oracle:
    spin-up/down radial functions of two species, random local tau00 with
    off-diagonal spin blocks, random moment directions; 1e-12 relative
    (green(r) to 1e-12 of its largest value)."""
    from tests.test_single_site_cpu import _usu
    rr, vr, jm, jw, rs, e, pr = CF.single_site_batch(nspecies=2, ne=3)
    matom, tmat, zlr, jlr, rc = ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr)
    assert rc == 0
    rng = np.random.default_rng(21)
    solves = [(ie, sp) for ie in range(3) for sp in range(2)] + [(1, 0)]
    n = len(solves)
    z = np.stack([np.stack([zlr[ie, 2 * sp], zlr[ie, 2 * sp + 1]]) for ie, sp in solves])
    j = np.stack([np.stack([jlr[ie, 2 * sp], jlr[ie, 2 * sp + 1]]) for ie, sp in solves])
    pot_of = np.array([2 * sp for _, sp in solves])
    tau = 0.2 * (rng.normal(size=(n, 4, 16, 16)) + 1j * rng.normal(size=(n, 4, 16, 16)))
    w = []
    for b in range(n):
        ev = rng.normal(size=3)
        w.append(np.stack(_usu(*CF.spin_rotation(ev / np.linalg.norm(ev)))))
    w = np.array(w)
    dos, green = ctx.green_mt_canted(3, tau, z, j, w, pot_of, rr, jm, jw, rs)
    for b in range(n):
        p = pot_of[b]
        # oracle layouts: tau (kkrsz, kkrsz, 4), radial (iprpts, lmax+1, 2)
        d_or, g_or = O.green_mt_canted(3, np.moveaxis(tau[b], 0, 2), np.transpose(z[b], (2, 1, 0)),
                                       np.transpose(j[b], (2, 1, 0)), rr[p], jm[p], jw[p], rs[p],
                                       w[b, 0], w[b, 1], w[b, 2])
        assert np.abs(dos[b] - d_or).max() <= 1e-12 * np.abs(d_or).max(), (b, dos[b], d_or)
        nj = jw[p]
        assert np.abs(green[b, :, :nj] - g_or).max() <= 1e-12 * np.abs(g_or).max()
        assert not np.any(green[b, :, nj:])


def test_store_filled_on_device_spin_canted():
    """Spin-canted store filled on the device (trltog of each species'
    spin-up and spin-down t) gives the tau00 of the same global-frame t built
    on the host (configs.local_to_global_t) and uploaded, and meets LAPACK."""
    import scipy.linalg as sla
    from tests import adjoint_ref as R
    e, w = CF.contour_semicircle(-0.5, 0.7, 6)
    s = CF.make_system("canted", "bcc", 5.42, (1, 1, 1), 6.0, 3, e, w, nspecies=2, seed=7,
                       n_spin_cant=2, n_spin_pola=2)
    pots = [CF.model_potential(26 + 3 * sp, 2.3, vmtz=-1.2 + 0.1 * isp)
            for sp in range(2) for isp in range(2)]
    rr = np.stack([p[0] for p in pots])
    vr = np.stack([p[1] for p in pots])
    jm = np.array([p[2] for p in pots])
    jw = np.array([p[3] for p in pots])
    rs = np.array([p[4] for p in pots])
    prels = np.array([s.prel(i) for i in range(len(s.energies))])
    uu = [CF.spin_rotation(s.evecs[sp]) for sp in range(2)]
    c = capi.Context(0)
    c.load_system(s)
    _, tmat, _, _, rc = c.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, s.energies, prels,
                                             radial=False)
    assert rc == 0
    c.tmats_from_single_site_canted(s.tmatStore.shape[0], s.tmatStore.shape[1],
                                    s.blkSizeTmatStore, len(s.energies), [0, 1], [0, 2], [1, 3],
                                    np.array([u for u, _ in uu]), np.array([ud for _, ud in uu]))
    store = np.zeros_like(s.tmatStore)
    for sp in range(2):
        for ie in range(len(s.energies)):
            tg = CF.local_to_global_t(tmat[ie, 2 * sp], tmat[ie, 2 * sp + 1], s.evecs[sp])
            base = ie * s.blkSizeTmatStore
            store[base:base + 1024, sp] = tg.flatten(order="F")
    s.tmatStore = store
    na = len(s.atoms)
    atoms = np.repeat(np.arange(na), 3)
    ies = np.tile([0, 2, 5], na)
    pr = [s.prel(i) for i in ies]
    tau_dev, info = c.tau00_batched(atoms, ies, pr, 32)
    c.close()
    c2 = capi.Context(0)
    c2.load_system(s)
    tau_host, info2 = c2.tau00_batched(atoms, ies, pr, 32)
    c2.close()
    assert np.all(info == 0) and np.all(info2 == 0)
    for b in range(len(atoms)):
        assert np.linalg.norm(tau_dev[b] - tau_host[b]) <= 1e-13 * np.linalg.norm(tau_host[b])
    for b in (0, len(atoms) - 1):
        m = R.oracle_m(s, atoms[b], ies[b])
        rhs = np.zeros((m.shape[0], 32), dtype=np.complex128)
        rhs[:32] = R.oracle_t0(s, atoms[b], ies[b])
        ref = sla.solve(m, rhs)[:32]
        assert np.linalg.norm(tau_dev[b] - ref) / np.linalg.norm(ref) < 1e-10


@pytest.mark.parametrize("lmax", [1, 2])
def test_lower_lmax_matches_oracle(ctx, lmax):
    """lmax = 1 and 2: matom, t, zlr, jlr and the muffin-tin DOS / green(r)
    against the oracle (the kernels' channel loops stop at lmax)."""
    rr, vr, jm, jw, rs, e, pr = CF.single_site_batch(nspecies=2, ne=4)
    nl, kk = lmax + 1, (lmax + 1) ** 2
    matom, tmat, zlr, jlr, rc = ctx.single_site_nonrel(0, lmax, vr, rr, jm, jw, rs, e, pr)
    assert rc == 0 and matom.shape == (4, len(rr), nl) and tmat.shape == (4, len(rr), kk, kk)
    for q in range(len(rr)):
        n = jw[q]
        for k in range(len(e)):
            m, t, z, j, _ = O.single_scatterer_nonrel(0, lmax, e[k], pr[k], vr[q], rr[q], jm[q],
                                                      jw[q], rs[q])
            assert np.max(np.abs(matom[k, q] - m) / np.abs(m)) < TOL
            assert np.abs(tmat[k, q] - t).max() <= TOL * np.abs(t).max()
            for l in range(nl):
                assert colerr(zlr[k, q, l, :n], z[:n, l]) < TOL
                assert colerr(jlr[k, q, l, :n], j[:n, l]) < TOL
    nsolve = len(rr) * len(e)
    pot_of = np.tile(np.arange(len(rr)), len(e))
    z = zlr.reshape(nsolve, nl, -1)
    j = jlr.reshape(nsolve, nl, -1)
    tau = tmat.reshape(nsolve, kk, kk)
    dos, green = ctx.green_mt(lmax, tau, z, j, pot_of, rr, jm, jw, rs)
    for b in range(0, nsolve, 3):
        p = pot_of[b]
        d_or, g_or, _, _ = O.green_mt(lmax, tau[b], z[b].T, j[b].T, rr[p], jm[p], jw[p], rs[p], 1)
        assert abs(dos[b] - d_or) <= 1e-12 * abs(d_or)
        assert np.abs(green[b, :jw[p]] - g_or).max() <= 1e-12 * np.abs(g_or).max()


def test_interstitial_matches_oracle(ctx):
    """The interstitial cell integrals (inter_m / interstitial.f) and their
    DOS on the GPU against the oracle, for matom of real potentials (GPU
    single-site solve), random tau00 near t, two cells (a cube of side 5.0
    around a muffin-tin radius 2.3, and a spherical cell reaching from inside
    the sphere), both zj conventions and both spin normalisations; and the
    whole-cell DOS (lsmsx_green_mt + lsmsx_green_interstitial) against the
    oracle's mdosms of pzzck + pzz and pzjck + pzj.  Bar 1e-12 relative."""
    rr, vr, jm, jw, rs, e, pr = CF.single_site_batch(nspecies=2, ne=3)
    npot, ne = len(rr), len(e)
    matom, tmat, zlr, jlr, rc = ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr)
    assert rc == 0
    nsolve = npot * ne
    pot_of = np.tile(np.arange(npot), ne)
    e_of = np.repeat(np.arange(ne), npot)
    rng = np.random.default_rng(17)
    ml = matom.reshape(nsolve, 4)
    mr = ml * (1.0 + 0.01 * rng.normal(size=ml.shape))  # left != right (spin-canted use)
    tau = tmat.reshape(nsolve, 16, 16) + 0.05 * (rng.normal(size=(nsolve, 16, 16)) +
                                                1j * rng.normal(size=(nsolve, 16, 16)))
    ctx.set_tables(3)  # the Gaunt numbers of the step-function sums
    cells = [CF.cube_cell(3, 5.0, 2.3), CF.sphere_cell(3, 2.0, 2.9, ngauss=8)]
    cell_of = np.arange(nsolve) % 2
    rsc = np.array([2.3, 2.3])
    prels = pr[e_of]
    for zj in (0, 1):
        for nspol in (1, 2):
            dos, pzz, pzj = ctx.green_interstitial(3, prels, ml, mr, tau, cells, cell_of, rsc,
                                                   zj_flag=zj, n_spin_pola=nspol)
            for b in range(nsolve):
                pz, pj = O.interstitial(zj, 3, prels[b], ml[b], mr[b], rsc[cell_of[b]],
                                        cells[cell_of[b]])
                assert np.abs(pzz[b] - pz).max() <= 1e-12 * np.abs(pz).max(), b
                if zj:
                    assert abs(pzj[b] - pj) <= 1e-12 * abs(pj), b
                else:
                    assert pzj[b] == 0
                d = -(2.0 / nspol) * (np.sum(pz * tau[b]) - pj) / np.pi
                assert abs(dos[b] - d) <= 1e-12 * abs(d), (b, dos[b], d)
    # whole cell = sphere part + interstitial part
    z = zlr.reshape(nsolve, 4, -1)
    j = jlr.reshape(nsolve, 4, -1)
    d_mt, _ = ctx.green_mt(3, tau, z, j, pot_of, rr, jm, jw, rs)
    r_cell = rs[pot_of]  # each solve's own muffin-tin radius
    cells1 = [CF.cube_cell(3, 5.0, float(r)) for r in rs]
    d_in, _, _ = ctx.green_interstitial(3, prels, ml, ml, tau, cells1, pot_of, rs)
    for b in range(nsolve):
        p = pot_of[b]
        d_or, _, _, _ = O.green_mt(3, tau[b], z[b].T, j[b].T, rr[p], jm[p], jw[p], rs[p], 1)
        pz, pj = O.interstitial(1, 3, prels[b], ml[b], ml[b], r_cell[b], cells1[p])
        whole = d_or - 2.0 * (np.sum(pz * tau[b]) - pj) / np.pi
        assert abs(d_mt[b] + d_in[b] - whole) <= 1e-12 * abs(whole), b
