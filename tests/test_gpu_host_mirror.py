"""The LSMS-shaped C++ host interface (lsms_amd/host/lsmsx_lsms.hpp), driven
from C++ exactly as calculateTauMatrix / calculateAllTauMatrices would call
it, against the oracle -- needs a GPU.

Checked: calculateAllTauMatricesMI355X output tau00_l = rotate(tau00 - t0)
for every atom and energy (collinear, spin-polarized collinear and
non-collinear systems), buildKKRMatrixMI355X, solveTau00MI355X and
calculateTauMatrixMI355X.  tau00 tolerance 1e-10 relative (plain FP64).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from lsms_amd import build as B
from lsms_amd import configs as CF
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dmat_of(a, k):
    """A seeded unitary spin rotation of atom a for the relativistic local
    frame (AtomData dmat; dmatp = dmat^dagger)."""
    rng = np.random.default_rng(100 + a)
    q, _ = np.linalg.qr(rng.normal(size=(k, k)) + 1j * rng.normal(size=(k, k)))
    return q, q.conj().T


def write_input(path, s, relativity=0):
    nsc, nsp = s.n_spin_cant, s.n_spin_pola
    ts = np.asfortranarray(s.tmatStore)
    with open(path, "wb") as f:
        np.array([s.maxlmax, nsc, nsp, len(s.atoms), ts.shape[1], len(s.energies),
                  s.blkSizeTmatStore, ts.shape[0], relativity], dtype=np.int32).tofile(f)
        ts.flatten(order="F").astype(np.complex128).tofile(f)
        s.energies.astype(np.complex128).tofile(f)
        for a, at in enumerate(s.atoms):
            np.array([at.numLIZ], dtype=np.int32).tofile(f)
            np.asfortranarray(at.LIZPos).flatten(order="F").tofile(f)
            at.LIZlmax.astype(np.int32).tofile(f)
            at.LIZStoreIdx.astype(np.int32).tofile(f)
            u, ud = O.spin_trafo(at.evec)
            u.astype(np.complex128).tofile(f)
            ud.astype(np.complex128).tofile(f)
            if relativity == 2:
                dm, dmp = dmat_of(a, s.kkrsz_ns)
                dm.flatten(order="F").astype(np.complex128).tofile(f)
                dmp.flatten(order="F").astype(np.complex128).tofile(f)


def oracle_tau00_l(s, a, ie, ispin):
    at = s.atoms[a]
    M = O.build_kkr(s.maxlmax, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore,
                    s.blkSizeTmatStore, ie, s.prel(ie), s.n_spin_cant, s.n_spin_pola, ispin)
    t0 = O.t0(s.maxlmax, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, ie,
              s.n_spin_cant, s.n_spin_pola, ispin)
    tau, _ = O.solve_tau00_zgetrf(M, t0)
    d = tau - t0
    if s.n_spin_cant == 2:
        u, ud = O.spin_trafo(at.evec)
        kk = (s.maxlmax + 1) ** 2
        return O.trgtol(kk, kk, u, ud, d).flatten(order="F"), tau, M
    return d.flatten(order="F"), tau, M


def run_driver(s, relativity=0):
    if not os.path.exists(B.DRIVER):
        pytest.fail("host_mirror_driver not built (python -m lsms_amd.build)")
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(fin, s, relativity)
        r = subprocess.run([B.DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        return np.fromfile(fout, dtype=np.complex128)


def systems():
    e, w = CF.contour_semicircle(-0.5, 0.7, 3)
    out = [("c1", CF.config_c1(npts=3))]
    out.append(("collinear", CF.make_system("col", "bcc", 5.42, (1, 1, 1), 6.0, 3, e, w, seed=5,
                                            n_spin_cant=1, n_spin_pola=2)))
    out.append(("noncollinear", CF.make_system("nc", "bcc", 5.42, (1, 1, 1), 6.0, 3, e, w, seed=6,
                                               n_spin_cant=2, n_spin_pola=2)))
    return out


@pytest.mark.parametrize("name,s", systems(), ids=[n for n, _ in systems()])
def test_host_mirror_matches_oracle(name, s):
    out = run_driver(s)
    na, ne = len(s.atoms), len(s.energies)
    nspin = 2 if s.n_spin_pola != s.n_spin_cant else 1
    k = s.kkrsz_ns
    pos = 0
    for ie in range(ne):
        blk = out[pos:pos + k * k * na * nspin].reshape((k * k, na * nspin), order="F")
        pos += k * k * na * nspin
        for sp in range(nspin):
            for a in range(na):
                ref, _, _ = oracle_tau00_l(s, a, ie, sp)
                got = blk[:, a + sp * na]
                assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref), (ie, a, sp)
    n = s.nrmat_ns(0)
    M = out[pos:pos + n * n].reshape((n, n), order="F")
    pos += n * n
    ref_l, ref_tau, ref_M = oracle_tau00_l(s, 0, 0, 0)
    assert np.abs(M - ref_M).max() <= 1e-12 * np.abs(ref_M).max()
    tau = out[pos:pos + k * k].reshape((k, k), order="F")
    pos += k * k
    assert np.linalg.norm(tau - ref_tau) <= 1e-10 * np.linalg.norm(ref_tau)
    tl = out[pos:pos + k * k]
    pos += k * k
    assert np.linalg.norm(tl - ref_l) <= 1e-10 * np.linalg.norm(ref_l)
    from tests import adjoint_ref as R
    tf = out[pos:pos + n * n].reshape((n, n), order="F")
    ref_f = R.tau_full(s, 0, 0)
    assert np.linalg.norm(tf - ref_f) <= 1e-10 * np.linalg.norm(ref_f)


def test_host_mirror_relativistic():
    """relativity = full: the kappa-mu store uploaded with its energies,
    calculateAllTauMatricesMI355X and calculateTauMatrixMI355X returning
    dmatp tau00 dmat (no tau00 - t; tau00Postprocess.cpp:185-192), and the
    kappa-mu M of buildKKRMatrixMI355X, against the oracle's relativistic
    build (relmtrx.f, setgij.f:39-55) with LAPACK."""
    s = CF.config_rel(npts=3)
    out = run_driver(s, relativity=2)
    na, ne, k = len(s.atoms), len(s.energies), s.kkrsz_ns
    blk = s.blkSizeTmatStore

    def ref(a, ie):
        at = s.atoms[a]
        m = O.build_kkr_rel(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore, blk, ie,
                            s.prel(ie), s.energies[ie])
        t0 = s.tmatStore[ie * blk:(ie + 1) * blk, at.LIZStoreIdx[0]].reshape(k, k, order="F")
        tau, _ = O.solve_tau00_zgetrf(m, t0)
        dm, dmp = dmat_of(a, k)
        return O.tripmt(dmp, tau, dm), tau, m

    pos = 0
    for ie in range(ne):
        blkout = out[pos:pos + k * k * na].reshape((k * k, na), order="F")
        pos += k * k * na
        for a in range(na):
            tl, _, _ = ref(a, ie)
            tl = tl.flatten(order="F")
            assert np.linalg.norm(blkout[:, a] - tl) <= 1e-10 * np.linalg.norm(tl), (ie, a)
    n = s.nrmat_ns(0)
    M = out[pos:pos + n * n].reshape((n, n), order="F")
    pos += n * n
    ref_l, ref_tau, ref_M = ref(0, 0)
    assert np.abs(M - ref_M).max() <= 1e-12 * np.abs(ref_M).max()
    tau = out[pos:pos + k * k].reshape((k, k), order="F")  # solveTau00MI355X on that M
    pos += k * k
    assert np.linalg.norm(tau - ref_tau) <= 1e-10 * np.linalg.norm(ref_tau)
    tl = out[pos:pos + k * k]
    pos += k * k
    assert np.linalg.norm(tl - ref_l.flatten(order="F")) <= 1e-10 * np.linalg.norm(ref_l)
    assert pos == len(out)


@pytest.mark.parametrize("nsc,nsp", [(1, 1), (1, 2), (2, 2)])
def test_single_site_host_mirror_fills_store(nsc, nsp):
    """solveSingleScatterersMI355X driven from C++ as energyContourIntegration
    would: every atom's slot of the t-matrix store equals the oracle's
    single-site t (single_scatterer_nonrel.f) laid out as
    solveSingleScatterers does -- spin blocks one after the other, or
    trltog(t_up, t_dn) for spin canting -- to 1e-10."""
    if not os.path.exists(B.SS_DRIVER):
        pytest.fail("single_site_driver not built (python -m lsms_amd.build)")
    e, _ = CF.contour_semicircle(-0.4, 0.7, 4)
    nat, lmax, kkr = 2, 3, 16
    blk = (2 * kkr) ** 2
    ldim = blk * len(e)
    rng = np.random.default_rng(4)
    atoms = []
    for i in range(nat):
        pots = [CF.model_potential(26 + 3 * i, 2.3, vmtz=-1.2 + 0.1 * s) for s in range(nsp)]
        ev = rng.normal(size=3)
        ev /= np.linalg.norm(ev)
        atoms.append((pots, ev))
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            iprpts = len(atoms[0][0][0][0])
            np.array([0, 0, nsc, nsp, nat, lmax, iprpts, len(e), blk, ldim],
                     dtype=np.int32).tofile(f)
            e.astype(np.complex128).tofile(f)
            for pots, ev in atoms:
                r, _, jmt, jws, rs = pots[0]
                r.astype(np.float64).tofile(f)
                np.array([jmt, jws], dtype=np.int32).tofile(f)
                np.array([rs, r[jws - 1]], dtype=np.float64).tofile(f)
                u, ud = O.spin_trafo(ev)
                u.astype(np.complex128).tofile(f)
                ud.astype(np.complex128).tofile(f)
                np.concatenate([p[1] for p in pots]).astype(np.float64).tofile(f)
        r = subprocess.run([B.SS_DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        store = np.fromfile(fout, dtype=np.complex128).reshape((ldim, nat), order="F")
    for i, (pots, ev) in enumerate(atoms):
        for ie, E in enumerate(e):
            pr = CF.prel_of(E)
            ts = [O.single_scatterer_nonrel(0, 3, E, pr, p[1], p[0], p[2], p[3], p[4])[1]
                  for p in pots]
            if nsc == 2:
                ref = CF.local_to_global_t(ts[0], ts[1], ev).flatten(order="F")
            else:
                ref = np.concatenate([t.flatten(order="F") for t in ts])
            got = store[ie * blk: ie * blk + len(ref), i]
            assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max(), (i, ie)
            assert not np.any(store[ie * blk + len(ref):(ie + 1) * blk, i])
