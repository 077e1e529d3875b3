"""Synthetic inputs for the tau00 path: lattices, local interaction zones
(LIZ), the energy contour and per-species t-matrix stores.

These mirror what the reference's setup code produces for the path:
  - LIZ: sphere cut around each local atom over periodic images, sorted by
    distance, central site first (Main/buildLIZandCommLists.cpp:136-190,
    377-400); per-site lmax reduced by `rsteps` (:393-399);
  - tmatStore: one column per t-matrix type, blocks of (2 kkrsz)^2 complex
    per energy slot (buildLIZandCommLists.cpp:553-555);
  - contour: Gauss-Legendre semicircle (Misc/congauss_c.f:3-124), without the
    zero-weight extra point the reference appends.
The t-matrices are synthetic: t_l = -sin(d_l) e^{i d_l} / kappa with seeded
phase shifts d_l per (species, l), diagonal in (l, m).
"""
from dataclasses import dataclass, field
import numpy as np

C_PHOT = 2.0 * 137.035999206  # include/PhysicalConstants.hpp:5-8


def prel_of(e):
    """kappa = sqrt(E (1 + E/c^2)) (calculateTauMatrix.cpp:846)."""
    e = np.asarray(e, dtype=np.complex128)
    return np.sqrt(e * (1.0 + e / (C_PHOT * C_PHOT)))


def dprel_of(e):
    """d kappa / dE = (1 + 2E/c^2) / (2 kappa)."""
    e = np.asarray(e, dtype=np.complex128)
    return (1.0 + 2.0 * e / (C_PHOT * C_PHOT)) / (2.0 * prel_of(e))


def contour_semicircle(ebot, etop, npts):
    """Gauss-Legendre semicircle contour, nodes ascending from ebot to etop."""
    x, w = np.polynomial.legendre.leggauss(npts)
    erad = 0.5 * (etop - ebot)
    zz = erad * np.exp(1j * 0.5 * np.pi * (1.0 - x))
    egrd = 0.5 * (etop + ebot) + zz
    dele = -1j * 0.5 * np.pi * zz * w
    return egrd.astype(np.complex128), dele.astype(np.complex128)


def lattice_sites(kind, a, ncell):
    """Cartesian sites of an n1 x n2 x n3 supercell of the conventional cubic
    cell; returns (positions (nat,3), box lengths (3,))."""
    if kind == "bcc":
        basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    elif kind == "fcc":
        basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    elif kind == "sc":
        basis = np.array([[0, 0, 0]])
    else:
        raise ValueError(kind)
    n1, n2, n3 = ncell
    pts = []
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                for b in basis:
                    pts.append((np.array([i, j, k]) + b) * a)
    return np.array(pts, dtype=np.float64), np.array([n1, n2, n3], dtype=np.float64) * a


def build_liz(pos, box, center, rcut, occupied=None):
    """Sites within rcut of pos[center] over periodic images, sorted by
    distance (ties broken by x, y, z), central site first.
    Returns (rel_pos (nliz,3), site_index (nliz,))."""
    if occupied is None:
        occupied = np.ones(len(pos), dtype=bool)
    c = pos[center]
    nimg = np.ceil(rcut / box).astype(int) + 1
    out = []
    r2 = rcut * rcut
    for i in range(-nimg[0], nimg[0] + 1):
        for j in range(-nimg[1], nimg[1] + 1):
            for k in range(-nimg[2], nimg[2] + 1):
                d = pos + np.array([i, j, k]) * box - c
                dd = np.einsum("ij,ij->i", d, d)
                sel = np.nonzero((dd <= r2) & occupied)[0]
                for s in sel:
                    out.append((dd[s], d[s, 0], d[s, 1], d[s, 2], s))
    out.sort(key=lambda t: (round(t[0], 9), round(t[1], 9), round(t[2], 9), round(t[3], 9)))
    rel = np.array([[t[1], t[2], t[3]] for t in out], dtype=np.float64)
    idx = np.array([t[4] for t in out], dtype=np.int64)
    assert np.allclose(rel[0], 0.0), "central site must come first"
    return rel, idx


def liz_lmax(rel, lmax, rsteps=None):
    """Per-site lmax: one step down per rsteps radius exceeded
    (buildLIZandCommLists.cpp:393-399)."""
    d = np.sqrt(np.sum(rel * rel, axis=1))
    out = np.full(len(rel), lmax, dtype=np.int32)
    if rsteps is not None:
        for rs in rsteps:
            out -= (d > rs).astype(np.int32)
    return out


def tmat_diag(deltas, kappa, lmax):
    """Diagonal t_lm = -sin(d_l) e^{i d_l} / kappa for one spin channel."""
    kkr = (lmax + 1) ** 2
    t = np.zeros(kkr, dtype=np.complex128)
    for l in range(lmax + 1):
        d = deltas[l]
        t[l * l:(l + 1) * (l + 1)] = -np.sin(d) * np.exp(1j * d) / kappa
    return t


@dataclass
class Atom:
    """Mirror of the AtomData fields the path reads (SingleSite/AtomData.hpp:441-467)."""
    numLIZ: int
    LIZPos: np.ndarray        # (3, numLIZ) column-major, relative to site 0
    LIZlmax: np.ndarray       # (numLIZ,) int32
    LIZStoreIdx: np.ndarray   # (numLIZ,) int32
    lmax: int
    kkrsz: int
    nrmat: int
    evec: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 1.0]))


@dataclass
class System:
    """A set of local atoms, the energy group and the t-matrix store."""
    name: str
    maxlmax: int
    n_spin_cant: int
    n_spin_pola: int
    atoms: list
    energies: np.ndarray      # (ne,) complex
    weights: np.ndarray       # (ne,) complex contour weights (zero where not on a contour)
    tmatStore: np.ndarray     # (lDim, num_store) complex, Fortran order
    blkSizeTmatStore: int
    deltas: np.ndarray        # (num_store, nspin, lmax+1) phase shifts
    evecs: np.ndarray = None  # (num_store, 3) moment directions (spin-canted systems)

    @property
    def kkrsz_ns(self):
        return self.n_spin_cant * (self.maxlmax + 1) ** 2

    def nrmat_ns(self, a):
        return self.n_spin_cant * self.atoms[a].nrmat

    def prel(self, ie):
        return complex(prel_of(self.energies[ie]))

    def dtmat_store(self):
        """dt/dE for every store entry.  The phase shifts do not depend on E,
        so every t_l carries 1/kappa and dt/dE = -(kappa'/kappa) t."""
        d = np.array(self.tmatStore, order="F", copy=True)
        blk = self.blkSizeTmatStore
        for ie, e in enumerate(self.energies):
            d[ie * blk:(ie + 1) * blk, :] *= -dprel_of(e) / prel_of(e)
        return d

    def at_energies(self, energies, weights=None):
        """The same atoms and t-matrix types with a new energy group."""
        energies = np.asarray(energies, dtype=np.complex128)
        store, blk = make_tmat_store(self.deltas, energies, self.maxlmax, self.n_spin_cant,
                                     self.n_spin_pola, self.evecs)
        w = np.zeros(len(energies), dtype=np.complex128) if weights is None else weights
        return System(name=self.name, maxlmax=self.maxlmax, n_spin_cant=self.n_spin_cant,
                      n_spin_pola=self.n_spin_pola, atoms=self.atoms, energies=energies,
                      weights=np.asarray(w, dtype=np.complex128), tmatStore=store,
                      blkSizeTmatStore=blk, deltas=self.deltas, evecs=self.evecs)


def make_tmat_store(deltas, energies, lmax, n_spin_cant, n_spin_pola, evecs_by_store=None):
    """tmatStore(iie*blk + i + j*kkrsz_ns, s) for every store column s and energy.
    Layout of one block: kkrsz_ns x kkrsz_ns with ld kkrsz_ns for
    n_spin_pola == n_spin_cant; collinear: the spin-up and spin-down
    kkrsz x kkrsz blocks one after the other (offset kkrsz^2 * ispin)."""
    kkr = (lmax + 1) ** 2
    blk = (2 * kkr) ** 2
    nstore = deltas.shape[0]
    ne = len(energies)
    store = np.zeros((blk * ne, nstore), dtype=np.complex128, order="F")
    for s in range(nstore):
        for ie, e in enumerate(energies):
            kap = complex(prel_of(e))
            base = ie * blk
            if n_spin_cant == 1 and n_spin_pola == 1:
                t = np.diag(tmat_diag(deltas[s, 0], kap, lmax))
                store[base:base + kkr * kkr, s] = t.flatten(order="F")
            elif n_spin_cant == 1 and n_spin_pola == 2:
                for isp in range(2):
                    t = np.diag(tmat_diag(deltas[s, isp], kap, lmax))
                    store[base + isp * kkr * kkr:base + (isp + 1) * kkr * kkr, s] = t.flatten(order="F")
            else:
                tu = np.diag(tmat_diag(deltas[s, 0], kap, lmax))
                td = np.diag(tmat_diag(deltas[s, 1], kap, lmax))
                ev = np.array([0.0, 0.0, 1.0]) if evecs_by_store is None else evecs_by_store[s]
                tg = local_to_global_t(tu, td, ev)
                store[base:base + 4 * kkr * kkr, s] = tg.flatten(order="F")
    return store, blk


def spin_rotation(evec):
    """u, u^dagger for a moment direction (Misc/spin_trafo.f)."""
    u1 = np.sqrt(0.5 * (1.0 + evec[2])) + 0j
    if abs(u1) < 1e-8:
        u = np.array([0.0, 1.0, 1.0, 0.0], dtype=np.complex128)
    else:
        u2 = -0.5 * (evec[0] + 1j * evec[1]) / u1
        u3 = 0.5 * (evec[0] - 1j * evec[1]) / u1
        u = np.array([u1, u2, u3, u1], dtype=np.complex128)
    ud = np.array([np.conj(u[0]), np.conj(u[2]), np.conj(u[1]), np.conj(u[3])])
    return u, ud


def local_to_global_t(tu, td, evec):
    """Global-frame 2kkr x 2kkr t from local spin-diagonal (tu, td):
    t_g = (U^dagger (x) 1) diag(tu, td) (U (x) 1) with U = [[u1,u3],[u2,u4]],
    the inverse of the local-frame rotation applied to tau00."""
    u, ud = spin_rotation(evec)
    U = np.array([[u[0], u[2]], [u[1], u[3]]])
    UD = np.array([[ud[0], ud[2]], [ud[1], ud[3]]])
    kkr = tu.shape[0]
    tl = np.zeros((2 * kkr, 2 * kkr), dtype=np.complex128)
    tl[:kkr, :kkr] = tu
    tl[kkr:, kkr:] = td
    I = np.eye(kkr)
    return np.kron(UD, I) @ tl @ np.kron(U, I)


def make_system(name, kind, a, ncell, rcut, lmax, energies, weights, nspecies=2,
                seed=20261202, local_atoms=None, n_spin_cant=1, n_spin_pola=1,
                vacancy=0.0, rcut_per_atom=None, rsteps=None, delta_range=1.2,
                _max_local=None):
    """Build a System from a lattice; species are drawn per site (seeded)."""
    rng = np.random.default_rng(seed)
    pos, box = lattice_sites(kind, a, ncell)
    nat = len(pos)
    if nat <= nspecies:
        species = np.arange(nat) % nspecies
    else:
        species = rng.integers(0, nspecies, nat)
    occupied = np.ones(nat, dtype=bool)
    if vacancy > 0:
        occupied = rng.random(nat) >= vacancy
    nspin = 2 if (n_spin_pola == 2) else 1
    deltas = rng.uniform(-delta_range, delta_range, size=(nspecies, nspin, lmax + 1))
    evecs = None
    if n_spin_cant == 2:
        v = rng.normal(size=(nspecies, 3))
        evecs = v / np.linalg.norm(v, axis=1, keepdims=True)
    if local_atoms is None:
        local_atoms = list(range(nat))
    local_atoms = [i for i in local_atoms if occupied[i]]
    if _max_local is not None:
        local_atoms = local_atoms[:_max_local]
    atoms = []
    for n, i in enumerate(local_atoms):
        rc = rcut if rcut_per_atom is None else rcut_per_atom[n % len(rcut_per_atom)]
        rel, idx = build_liz(pos, box, i, rc, occupied)
        lm = liz_lmax(rel, lmax, rsteps)
        ev = np.array([0.0, 0.0, 1.0]) if evecs is None else evecs[species[i]]
        atoms.append(Atom(numLIZ=len(rel), LIZPos=np.asfortranarray(rel.T),
                          LIZlmax=lm.astype(np.int32), LIZStoreIdx=species[idx].astype(np.int32),
                          lmax=lmax, kkrsz=(lmax + 1) ** 2,
                          nrmat=int(np.sum((lm + 1) ** 2)), evec=ev))
    store, blk = make_tmat_store(deltas, energies, lmax, n_spin_cant, n_spin_pola, evecs)
    return System(name=name, maxlmax=lmax, n_spin_cant=n_spin_cant, n_spin_pola=n_spin_pola,
                  atoms=atoms, energies=np.asarray(energies, dtype=np.complex128),
                  weights=np.asarray(weights, dtype=np.complex128), tmatStore=store,
                  blkSizeTmatStore=blk, deltas=deltas, evecs=evecs)


def config_c1(npts=31, seed=20261202):
    """bcc a=5.42 bohr, 2-atom cubic cell, LIZ r=6.0 (15 sites, N=240), lmax=3,
    31-point semicircle contour -0.5..0.7 Ry."""
    e, w = contour_semicircle(-0.5, 0.7, npts)
    return make_system("c1_bcc15", "bcc", 5.42, (1, 1, 1), 6.0, 3, e, w, nspecies=2, seed=seed)


def config_c2(npts=31, seed=20261202, natoms=32):
    """fcc a=6.82 bohr, 2x2x2 conventional supercell (32 atoms), LIZ r=13.0
    (135 sites, N=2160), lmax=3, 31-point semicircle contour."""
    e, w = contour_semicircle(-0.5, 0.7, npts)
    return make_system("c2_fcc135", "fcc", 6.82, (2, 2, 2), 13.0, 3, e, w, nspecies=2,
                       seed=seed, local_atoms=list(range(natoms)))


def config_c3_energies():
    """32 semicircle points plus 16 near the real axis: Re E in [-0.4, 0.7],
    Im E log-spaced 0.001..0.01 Ry (weights zero for the latter)."""
    e, w = contour_semicircle(-0.5, 0.7, 32)
    re = np.linspace(-0.4, 0.7, 16)
    im = np.logspace(-3, -2, 16)
    e2 = re + 1j * im
    return np.concatenate([e, e2]), np.concatenate([w, np.zeros(16, dtype=np.complex128)])


def config_c3(seed=20261202, natoms=32):
    e, w = config_c3_energies()
    return make_system("c3_fcc135_mixed", "fcc", 6.82, (2, 2, 2), 13.0, 3, e, w, nspecies=2,
                       seed=seed, local_atoms=list(range(natoms)))


def config_c5(npts=31, seed=20261202, natoms=8, rcut=13.3):
    """Spin-canted bcc a=5.42 bohr, n_spin_cant = n_spin_pola = 2, LIZ r=13.3
    (137 sites, N = 137*32 = 4384), lmax=3, 31-point semicircle contour.
    2x2x2 conventional cells (16 sites); every site is its own t-matrix type
    (store column = site) with a moment direction uniform on the sphere and
    the spin-up/down phase shifts of its species."""
    rng = np.random.default_rng(seed)
    lmax = 3
    e, w = contour_semicircle(-0.5, 0.7, npts)
    pos, box = lattice_sites("bcc", 5.42, (2, 2, 2))
    nat = len(pos)
    species = rng.integers(0, 2, nat)
    dspec = rng.uniform(-1.2, 1.2, size=(2, 2, lmax + 1))
    v = rng.normal(size=(nat, 3))
    evecs = v / np.linalg.norm(v, axis=1, keepdims=True)
    deltas = dspec[species]
    atoms = []
    for i in range(natoms):
        rel, idx = build_liz(pos, box, i, rcut)
        lm = liz_lmax(rel, lmax)
        atoms.append(Atom(numLIZ=len(rel), LIZPos=np.asfortranarray(rel.T),
                          LIZlmax=lm.astype(np.int32), LIZStoreIdx=np.asarray(idx, dtype=np.int32),
                          lmax=lmax, kkrsz=(lmax + 1) ** 2, nrmat=int(np.sum((lm + 1) ** 2)),
                          evec=evecs[i]))
    store, blk = make_tmat_store(deltas, e, lmax, 2, 2, evecs)
    return System(name="c5_bcc137_canted", maxlmax=lmax, n_spin_cant=2, n_spin_pola=2,
                  atoms=atoms, energies=e, weights=w, tmatStore=store, blkSizeTmatStore=blk,
                  deltas=deltas, evecs=evecs)


def config_c4(npts=32, seed=20261202, natoms=64, rcuts=(11.5, 13.3)):
    """Random binary bcc alloy, a=5.42 bohr, 16x16x8 conventional cells (4096
    sites, periodic), 10% vacancies (seeded), lmax=3; the first `natoms`
    occupied sites are local and each draws its LIZ radius from `rcuts`
    (seeded), so the KKR sizes are uneven; 32-point semicircle contour."""
    rng = np.random.default_rng(seed + 4)
    r_per_atom = list(rng.choice(np.asarray(rcuts, dtype=float), size=natoms))
    e, w = contour_semicircle(-0.5, 0.7, npts)
    # draw enough local candidates that `natoms` occupied ones remain
    return make_system("c4_bcc4096_alloy", "bcc", 5.42, (16, 16, 8), max(rcuts), 3, e, w,
                       nspecies=2, seed=seed, local_atoms=list(range(2 * natoms)),
                       vacancy=0.10, rcut_per_atom=r_per_atom, _max_local=natoms)


def log_mesh(rmt, jmt, h=0.02, iprpts=1000):
    """Logarithmic radial mesh r_j = exp(x0 + (j-1) h) with r_jmt = rmt
    (1-based jmt), the form of AtomData::r_mesh (xstart, h, jmt)."""
    x0 = np.log(rmt) - (jmt - 1) * h
    return np.exp(x0 + h * np.arange(iprpts))


def model_potential(z, rmt, jmt=450, jws=470, h=0.02, iprpts=1000, screen=0.6, vmtz=-1.2,
                    shift=0.0):
    """Synthetic spherical potential r v(r) = -2Z exp(-r/screen) + vmtz r + shift r^2
    (Rydberg units) on log_mesh(rmt, jmt): nuclear -2Z at the origin, screened
    outwards.  Returns (r_mesh, vr, jmt, jws, r_sph) with r_sph a little inside
    the muffin-tin radius and between mesh points, as rInscribed usually is."""
    r = log_mesh(rmt, jmt, h, iprpts)
    vr = -2.0 * z * np.exp(-r / screen) + vmtz * r + shift * r * r
    r_sph = 0.5 * (r[jmt - 2] + r[jmt - 1]) * 0.999
    return r, vr, jmt, jws, r_sph


def single_site_batch(nspecies=4, ne=48, seed=20261202, ebot=-0.4, etop=0.75):
    """Single-site workload of the C3 shape: nspecies x 2 spins potentials at
    ne energies on the C3 contour band (Z 22..29, rmt 2.2-2.5 bohr).
    Returns (r_mesh[npot, iprpts], vr[npot, iprpts], jmt[npot], jws[npot],
    r_sph[npot], energies[ne], prels[ne])."""
    rng = np.random.default_rng(seed)
    rs, vs, jm, jw, rsph = [], [], [], [], []
    for s in range(nspecies):
        z = 22 + int(rng.integers(0, 8))
        rmt = 2.2 + 0.3 * float(rng.random())
        for spin in (0, 1):
            r, vr, jmt, jws, r_sph = model_potential(z, rmt, screen=0.55 + 0.1 * float(rng.random()),
                                                      vmtz=-1.2 + 0.05 * (2 * spin - 1))
            rs.append(r)
            vs.append(vr)
            jm.append(jmt)
            jw.append(jws)
            rsph.append(r_sph)
    e, _ = contour_semicircle(ebot, etop, ne)
    return (np.array(rs), np.array(vs), np.array(jm, dtype=np.int32),
            np.array(jw, dtype=np.int32), np.array(rsph), e, prel_of(e))


def make_rel_tmat_store(nstore, energies, lmax, seed=20261203, delta_range=1.2, coupling=0.05):
    """Synthetic relativistic (kappa-mu) t-matrix store: per store column,
    phase shifts per (l, j) with t_kappa-mu = -sin(d) e^{i d} / kappa on the
    diagonal (within each l the 2l rows of j = l - 1/2, then the 2l + 2 rows
    of j = l + 1/2, the clebsch.f row order) plus a fixed complex symmetric
    coupling of size `coupling` (as spin-orbit and exchange mix kappa-mu
    states), all over kappa = prel(E).  Blocks of (2 kkr)^2 per energy."""
    rng = np.random.default_rng(seed)
    kkr = (lmax + 1) ** 2
    kk2 = 2 * kkr
    blk = kk2 * kk2
    deltas = rng.uniform(-delta_range, delta_range, size=(nstore, lmax + 1, 2))
    cpl = coupling * (rng.normal(size=(nstore, kk2, kk2)) + 1j * rng.normal(size=(nstore, kk2, kk2)))
    cpl = 0.5 * (cpl + np.transpose(cpl, (0, 2, 1)))
    ne = len(energies)
    store = np.zeros((blk * ne, nstore), dtype=np.complex128, order="F")
    for s in range(nstore):
        dvec = []
        for l in range(lmax + 1):
            dvec += [deltas[s, l, 0]] * (2 * l) + [deltas[s, l, 1]] * (2 * l + 2)
        dvec = np.array(dvec)
        base_t = np.diag(-np.sin(dvec) * np.exp(1j * dvec)) + cpl[s]
        for ie, e in enumerate(energies):
            t = base_t / complex(prel_of(e))
            store[ie * blk:(ie + 1) * blk, s] = t.flatten(order="F")
    return store, blk


def config_rel(npts=4, seed=20261203, rcut=6.0, ncell=(1, 1, 1)):
    """Relativistic test system (relativity = full): bcc a=5.42 bohr,
    n_spin_cant = n_spin_pola = 2, lmax 3 on every LIZ site (r=6.0: 15
    sites, N = 480), synthetic kappa-mu t-matrices per species on a
    semicircle contour."""
    e, w = contour_semicircle(-0.5, 0.7, npts)
    base = make_system("rel_bcc", "bcc", 5.42, ncell, rcut, 3, e, w, nspecies=2, seed=seed,
                       n_spin_cant=2, n_spin_pola=2)
    store, blk = make_rel_tmat_store(base.tmatStore.shape[1], e, 3, seed=seed)
    return System(name="rel_bcc", maxlmax=3, n_spin_cant=2, n_spin_pola=2, atoms=base.atoms,
                  energies=base.energies, weights=base.weights, tmatStore=store,
                  blkSizeTmatStore=blk, deltas=base.deltas, evecs=base.evecs)


def _gl(n, a, b):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def sphere_cell(lmax, r_lo, r_hi, ngauss=12):
    """Cell data (ncrit, ngaussr, grwylm, gwwylm, wylm) of a spherical cell of
    radius r_hi, integrated from r_lo (which may lie inside the muffin-tin
    sphere): one region, Gauss-Legendre points, weights times r^2 (the cell
    integral's Jacobian), step function sigma_00 = sqrt(4 pi), other
    sigma_lm = 0."""
    nw = (2 * lmax + 1) * (lmax + 1)
    r, w = _gl(ngauss, r_lo, r_hi)
    gr = r.reshape(ngauss, 1)
    gw = (w * r * r).reshape(ngauss, 1)
    wy = np.zeros((nw, ngauss, 1), dtype=np.complex128)
    wy[0] = np.sqrt(4.0 * np.pi)
    return 2, ngauss, np.asfortranarray(gr), np.asfortranarray(gw), np.asfortranarray(wy)


def cube_cell(lmax, a, r_sph, ngauss=8, nth=96, nph=192):
    """Cell data of a cube of side a (the simple-cubic Voronoi cell) outside
    the muffin-tin radius r_sph <= a/2: regions between r_sph and the
    critical radii h, h sqrt 2, h sqrt 3 (h = a/2), Gauss-Legendre points with
    weights times r^2, and the step function's expansion
    sigma_lm(r) = int dOmega Theta(r rhat) conj(Y_lm(rhat)) for m >= 0 (index
    l(l+1)/2 + m, l <= 2 lmax) by Gauss-Legendre in cos(theta) and a uniform
    phi grid (synthetic cell data of the shape VORPOL passes on)."""
    from scipy.special import sph_harm_y
    h = 0.5 * a
    crit = [r_sph, h, h * np.sqrt(2.0), h * np.sqrt(3.0)]
    ncrit = len(crit)
    lw = 2 * lmax
    nw = (lw + 1) * (lmax + 1)
    ct, wt = np.polynomial.legendre.leggauss(nth)
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    th = np.arccos(ct)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    W = np.outer(wt, np.full(nph, 2 * np.pi / nph))
    dirs = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], axis=-1)
    ylm = {}
    for l in range(lw + 1):
        for m in range(0, l + 1):
            ylm[(l, m)] = np.conj(sph_harm_y(l, m, TH, PH))
    gr = np.zeros((ngauss, ncrit - 1))
    gw = np.zeros((ngauss, ncrit - 1))
    wy = np.zeros((nw, ngauss, ncrit - 1), dtype=np.complex128)
    for ir in range(ncrit - 1):
        r, w = _gl(ngauss, crit[ir], crit[ir + 1])
        gr[:, ir] = r
        gw[:, ir] = w * r * r
        for ng in range(ngauss):
            inside = np.all(np.abs(r[ng] * dirs) <= h, axis=-1).astype(float)
            for l in range(lw + 1):
                for m in range(0, l + 1):
                    wy[l * (l + 1) // 2 + m, ng, ir] = np.sum(W * inside * ylm[(l, m)])
    return ncrit, ngauss, np.asfortranarray(gr), np.asfortranarray(gw), np.asfortranarray(wy)
