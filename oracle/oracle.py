"""ctypes bindings to the CPU oracle (oracle/liblsms_oracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP path: only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  The product package (lsms_amd) never imports it.

Each wrapper names the reference routine its C function restates; the C
source carries the file:line citations.
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblsms_oracle.so")
_lib = None

_zp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="F_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")


class _Z(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def _z(v):
    v = complex(v)
    return _Z(v.real, v.imag)


def lib():
    """Load the oracle library (build it with `make -C oracle`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
    L = C.CDLL(_LIB_PATH)
    L.or_prel.restype = _Z
    L.or_prel.argtypes = [_Z]
    L.or_gauss_legendre.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp]
    L.or_plm_normalized.argtypes = [C.c_int, C.c_double, _dp]
    L.or_cgaunt.argtypes = [C.c_int, _dp]
    L.or_ifacts.argtypes = [C.c_int, _zp, _zp]
    L.or_hankel.argtypes = [_Z, C.c_double, C.c_int, _zp]
    L.or_makegij.argtypes = [C.c_int, C.c_int, C.c_int, _Z, _dp, _dp, _zp, _zp, _zp]
    L.or_build_kkr.restype = C.c_int
    L.or_build_kkr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip,
                               _zp, C.c_long, C.c_long, C.c_int, _Z, _zp, C.c_int]
    L.or_t0.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _zp, C.c_long,
                        C.c_long, C.c_int, _zp]
    L.or_zgetrf.restype = C.c_int
    L.or_zgetrf.argtypes = [C.c_int, C.c_int, _zp, C.c_int, _ip]
    L.or_zgetrs.argtypes = [C.c_int, C.c_int, _zp, C.c_int, _ip, _zp, C.c_int]
    L.or_solve_tau00_zgetrf.restype = C.c_int
    L.or_solve_tau00_zgetrf.argtypes = [C.c_int, C.c_int, _zp, C.c_int, _zp, _zp]
    L.or_solve_tau00_zblocklu.restype = C.c_int
    L.or_solve_tau00_zblocklu.argtypes = [C.c_int, C.c_int, C.c_int, _zp, C.c_int, _zp, _zp]
    L.or_tau00_minus_t.argtypes = [C.c_int, _zp, _zp, _zp]
    L.or_spin_trafo.argtypes = [_dp, _zp, _zp]
    L.or_trgtol.argtypes = [C.c_int, C.c_int, _zp, _zp, _zp, _zp]
    L.or_congauss.argtypes = [C.c_double, C.c_double, C.c_int, _zp, _zp]
    L.or_single_scatterer_nonrel.restype = C.c_int
    L.or_single_scatterer_nonrel.argtypes = [C.c_int, C.c_int, _Z, _Z, _dp, _dp, C.c_int,
                                             C.c_int, C.c_double, C.c_int, _zp, _zp, _zp, _zp]
    L.or_fitpot.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int]
    L.or_sph_bessel_jh.argtypes = [C.c_int, _Z, _zp, _zp, _zp]
    L.or_cnewint.argtypes = [C.c_int, _dp, _zp, _zp, C.c_int]
    L.or_green_mt_canted.restype = C.c_int
    L.or_green_mt_canted.argtypes = [C.c_int, _zp, _zp, _zp, _dp, C.c_int, C.c_int, C.c_double,
                                     C.c_int, _zp, _zp, _zp]
    L.or_green_mt.restype = C.c_int
    L.or_green_mt.argtypes = [C.c_int, _zp, _zp, _zp, _dp, C.c_int, C.c_int, C.c_double, C.c_int,
                              C.c_int, _zp, _zp, _zp, _zp]
    L.or_interstitial.argtypes = [C.c_int, C.c_int, _Z, _zp, _zp, C.c_double, C.c_int, C.c_int,
                                  _dp, _dp, _zp, _dp, C.c_int, _zp, _zp]
    L.or_clebsch.argtypes = [_dp, _dp, _ip, _ip]
    L.or_relmtrx.argtypes = [_zp, C.c_int, _zp]
    L.or_build_kkr_rel.restype = C.c_int
    L.or_build_kkr_rel.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _zp, C.c_long, C.c_long,
                                   C.c_int, _Z, _Z, _zp, C.c_int]
    L.or_tripmt.argtypes = [_zp, _zp, _zp, C.c_int]
    _lib = L
    return L


def _f(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def prel(energy):
    """kappa = sqrt(E(1+E/c^2)) (calculateTauMatrix.cpp:846)."""
    r = lib().or_prel(_z(energy))
    return complex(r.re, r.im)


def gauss_legendre(n, x1=-1.0, x2=1.0):
    x = np.zeros(n, order="F")
    w = np.zeros(n, order="F")
    lib().or_gauss_legendre(n, x1, x2, x, w)
    return x, w


def plm_normalized(lmax, x):
    p = np.zeros((lmax + 1) * (lmax + 2) // 2, order="F")
    lib().or_plm_normalized(lmax, float(x), p)
    return p


def cgaunt(lmax):
    """cgnt(l3/2, j1, j2) as an array of shape (lmax+1, kkr, kkr)."""
    kkr = (lmax + 1) ** 2
    c = np.zeros((lmax + 1, kkr, kkr), order="F")
    lib().or_cgaunt(lmax, c)
    return c


def ifacts(lmax):
    kkr = (lmax + 1) ** 2
    illp = np.zeros((kkr, kkr), dtype=np.complex128, order="F")
    ilp1 = np.zeros(2 * lmax + 1, dtype=np.complex128, order="F")
    lib().or_ifacts(lmax, illp, ilp1)
    return illp, ilp1


def hankel(prel_, r, lend):
    h = np.zeros(max(lend, 1) + 1, dtype=np.complex128, order="F")
    lib().or_hankel(_z(prel_), float(r), lend, h)
    return h[: lend + 1]


def makegij(lmaxi, lmaxj, lmax_tab, prel_, rij, cgnt=None, illp=None, ilp1=None):
    if cgnt is None:
        cgnt = cgaunt(lmax_tab)
    if illp is None:
        illp, ilp1 = ifacts(lmax_tab)
    g = np.zeros(((lmaxi + 1) ** 2, (lmaxj + 1) ** 2), dtype=np.complex128, order="F")
    lib().or_makegij(lmaxi, lmaxj, lmax_tab, _z(prel_), _f(rij, np.float64),
                     _f(cgnt, np.float64), _f(illp, np.complex128),
                     _f(ilp1, np.complex128), g)
    return g


def build_kkr(maxlmax, liz_pos, liz_lmax, liz_store, tmat_store, blk_size, iie, prel_,
              n_spin_cant=1, n_spin_pola=1, ispin=0):
    """KKR matrix M = 1 - t G, F77 route (calculateTauMatrix.cpp:57-315)."""
    liz_pos = _f(liz_pos, np.float64)
    liz_lmax = _f(liz_lmax, np.int32)
    liz_store = _f(liz_store, np.int32)
    ts = _f(tmat_store, np.complex128)
    n = n_spin_cant * int(np.sum((liz_lmax + 1) ** 2))
    m = np.zeros((n, n), dtype=np.complex128, order="F")
    lib().or_build_kkr(maxlmax, n_spin_cant, n_spin_pola, ispin, liz_pos.shape[1], liz_pos,
                       liz_lmax, liz_store, ts, ts.shape[0], blk_size, iie, _z(prel_), m, n)
    return m


def t0(maxlmax, store0, tmat_store, blk_size, iie, n_spin_cant=1, n_spin_pola=1, ispin=0):
    k = n_spin_cant * (maxlmax + 1) ** 2
    ts = _f(tmat_store, np.complex128)
    out = np.zeros((k, k), dtype=np.complex128, order="F")
    lib().or_t0(maxlmax, n_spin_cant, n_spin_pola, ispin, store0, ts, ts.shape[0], blk_size,
                iie, out)
    return out


def zgetrf(a):
    a = np.array(a, dtype=np.complex128, order="F", copy=True)
    m, n = a.shape
    ipiv = np.zeros(min(m, n), dtype=np.int32, order="F")
    info = lib().or_zgetrf(m, n, a, m, ipiv)
    return a, ipiv, info


def solve_tau00_zgetrf(m, t0_):
    """tau00 = [M^-1 [t0;0]]_00 (linearSolvers_CPU.cpp:98-123)."""
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    k = t0_.shape[0]
    tau = np.zeros((k, k), dtype=np.complex128, order="F")
    info = lib().or_solve_tau00_zgetrf(a.shape[0], k, a, a.shape[0], _f(t0_, np.complex128), tau)
    return tau, info


def solve_tau00_zblocklu(m, t0_, nblk=3):
    """Block-inverse route (zblock_lu_cpp.cpp, linearSolvers_CPU.cpp:261-331)."""
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    k = t0_.shape[0]
    tau = np.zeros((k, k), dtype=np.complex128, order="F")
    info = lib().or_solve_tau00_zblocklu(a.shape[0], k, nblk, a, a.shape[0],
                                         _f(t0_, np.complex128), tau)
    return tau, info


def spin_trafo(evec):
    u = np.zeros(4, dtype=np.complex128, order="F")
    ud = np.zeros(4, dtype=np.complex128, order="F")
    lib().or_spin_trafo(_f(evec, np.float64), u, ud)
    return u, ud


def trgtol(nr, nc, u, ud, mg):
    ml = np.zeros((nr, nc, 4), dtype=np.complex128, order="F")
    lib().or_trgtol(nr, nc, _f(u, np.complex128), _f(ud, np.complex128),
                    _f(mg, np.complex128), ml)
    return ml


def congauss(ebot, etop, ngauss):
    e = np.zeros(ngauss, dtype=np.complex128, order="F")
    w = np.zeros(ngauss, dtype=np.complex128, order="F")
    lib().or_congauss(ebot, etop, ngauss, e, w)
    return e, w


def single_scatterer_nonrel(nrelv, lmax, energy, prel_, vr, r, jmt, jws, r_sph):
    """t-matrix solve of one potential at one energy
    (SingleSite/single_scatterer_nonrel.f:4-122 via semrel_c.f, solver 2).
    vr = r v(r) and r on the radial mesh (length iprpts); jmt, jws 1-based
    grid indices as in AtomData.  Returns (matom[lmax+1], tmat[kkrsz,kkrsz],
    zlr[iprpts,lmax+1], jlr[iprpts,lmax+1], rc)."""
    vr = _f(vr, np.float64)
    r = _f(r, np.float64)
    iprpts = len(r)
    kkrsz = (lmax + 1) ** 2
    matom = np.zeros(lmax + 1, dtype=np.complex128)
    tmat = np.zeros((kkrsz, kkrsz), dtype=np.complex128, order="F")
    zlr = np.zeros((iprpts, lmax + 1), dtype=np.complex128, order="F")
    jlr = np.zeros((iprpts, lmax + 1), dtype=np.complex128, order="F")
    rc = lib().or_single_scatterer_nonrel(nrelv, lmax, _z(energy), _z(prel_), vr, r, jmt, jws,
                                          r_sph, iprpts, matom, tmat, zlr, jlr)
    if rc < 0:
        raise ValueError("or_single_scatterer_nonrel: bad sizes")
    return matom, tmat, zlr, jlr, rc


def fitpot(r, rv, nfit, iend):
    """Misc/fitpot.f with nr = 4: coefficients (4, iend)."""
    coef = np.zeros((4, iend), dtype=np.float64, order="F")
    lib().or_fitpot(_f(r, np.float64), _f(rv, np.float64), coef, nfit, iend)
    return coef


def sph_bessel_jh(lmax, x):
    """Misc/zsphbesjh.f (job = 1): j_l(x), h_l(x) x e^{-ix}, e^{ix}."""
    bj = np.zeros(lmax + 1, dtype=np.complex128)
    hm = np.zeros(lmax + 1, dtype=np.complex128)
    eiz = np.zeros(1, dtype=np.complex128)
    lib().or_sph_bessel_jh(lmax, _z(x), bj, hm, eiz)
    return bj, hm, eiz[0]


def cnewint(x, f, ip0):
    """Misc/cnewint.f: running integral of f x^ip0 on the mesh x."""
    x = _f(x, np.float64)
    f = _f(f, np.complex128)
    g = np.zeros(len(x), dtype=np.complex128)
    lib().or_cnewint(len(x), x, f, g, ip0)
    return g


def green_mt(lmax, tau00, zlr, jlr, r_mesh, jmt, jws, r_sph, n_spin_pola=1):
    """Muffin-tin / ASA part of green_function.f for one collinear spin:
    (dos, green[jws], zlzl[lmax+1], zljl[lmax+1])."""
    tau00 = _f(tau00, np.complex128)
    zlr = _f(zlr, np.complex128)
    jlr = _f(jlr, np.complex128)
    r_mesh = _f(r_mesh, np.float64)
    iprpts = zlr.shape[0]
    dos = np.zeros(1, dtype=np.complex128)
    green = np.zeros(max(jws, 1), dtype=np.complex128)
    zz = np.zeros(lmax + 1, dtype=np.complex128)
    zj = np.zeros(lmax + 1, dtype=np.complex128)
    rc = lib().or_green_mt(lmax, tau00, zlr, jlr, r_mesh, jmt, jws, r_sph, iprpts, n_spin_pola,
                           dos, green, zz, zj)
    if rc != 0:
        raise ValueError("or_green_mt: bad sizes")
    return dos[0], green[:jws], zz, zj


def green_mt_canted(lmax, tau00_l, zlr, jlr, r_mesh, jmt, jws, r_sph, wx, wy, wz):
    """Spin-canted muffin-tin part of green_function.f: tau00_l (kkrsz, kkrsz, 4)
    column-major blocks, zlr/jlr (iprpts, lmax+1, 2).  Returns (dos[4], green[4, jws])."""
    tau00_l = _f(tau00_l, np.complex128)
    zlr = _f(zlr, np.complex128)
    jlr = _f(jlr, np.complex128)
    w = np.concatenate([np.asarray(wx), np.asarray(wy), np.asarray(wz)]).astype(np.complex128)
    dos = np.zeros(4, dtype=np.complex128)
    green = np.zeros(4 * max(jws, 1), dtype=np.complex128)
    rc = lib().or_green_mt_canted(lmax, tau00_l, zlr, jlr, _f(r_mesh, np.float64), jmt, jws,
                                  r_sph, zlr.shape[0], w, dos, green)
    if rc != 0:
        raise ValueError("or_green_mt_canted: bad sizes")
    return dos, green[: 4 * jws].reshape(4, jws)


# ---- relativistic KKR blocks -------------------------------------------------

def clebsch():
    """(u1, u2, ind1, ind2) of Misc/clebsch.f, 500 entries each (ind 1-based)."""
    u1 = np.zeros(500, order="F")
    u2 = np.zeros(500, order="F")
    i1 = np.zeros(500, dtype=np.int32, order="F")
    i2 = np.zeros(500, dtype=np.int32, order="F")
    lib().or_clebsch(u1, u2, i1, i2)
    return u1, u2, i1, i2


def clebsch_matrix(lmax):
    """U (2kkr x 2kkr): row = kappa-mu index, column = s*kkr + lm (s = 0 for
    the u1, s = 1 for the u2 component), so relmtrx(a) = U diag_spin(a) U^T."""
    kkr = (lmax + 1) ** 2
    u1, u2, i1, i2 = clebsch()
    U = np.zeros((2 * kkr, 2 * kkr))
    for i in range(2 * kkr):
        U[i, i1[i] - 1] += u1[i]
        U[i, kkr + i2[i] - 1] += u2[i]
    return U


def relmtrx(a):
    kkr = a.shape[0]
    b = np.zeros((2 * kkr, 2 * kkr), dtype=np.complex128, order="F")
    lib().or_relmtrx(_f(a, np.complex128), kkr, b)
    return b


def build_kkr_rel(maxlmax, liz_pos, liz_lmax, liz_store, tmat_store, blk_size, iie, prel_,
                  energy):
    """Relativistic M = 1 - t (prel^2/E) relmtrx(G) (setgij.f:39-55)."""
    liz_pos = _f(liz_pos, np.float64)
    liz_lmax = _f(liz_lmax, np.int32)
    liz_store = _f(liz_store, np.int32)
    ts = _f(tmat_store, np.complex128)
    n = 2 * (maxlmax + 1) ** 2 * liz_pos.shape[1]
    m = np.zeros((n, n), dtype=np.complex128, order="F")
    rc = lib().or_build_kkr_rel(maxlmax, liz_pos.shape[1], liz_pos, liz_lmax, liz_store, ts,
                                ts.shape[0], blk_size, iie, _z(prel_), _z(energy), m, n)
    if rc < 0:
        raise ValueError("relativistic build needs every LIZ site at maxlmax")
    return m


def tripmt(u, b, ust):
    out = np.array(b, dtype=np.complex128, order="F", copy=True)
    lib().or_tripmt(_f(u, np.complex128), out, _f(ust, np.complex128), out.shape[0])
    return out


def interstitial(zj_flag, lmax, pnrel, matom_left, matom_right, r_sph, cell, cgnt=None):
    """Interstitial cell integrals (VORPOL/interstitial.f): (pzz, pzj).
    cell = (ncrit, ngaussr, grwylm (ngaussr, ncrit-1), gwwylm, wylm (nw, ngaussr, ncrit-1))."""
    ncrit, ngaussr, gr, gw, wy = cell
    if cgnt is None:
        cgnt = cgaunt(lmax)
    kkr = (lmax + 1) ** 2
    pzz = np.zeros((kkr, kkr), dtype=np.complex128, order="F")
    pzj = np.zeros(1, dtype=np.complex128)
    lib().or_interstitial(int(zj_flag), lmax, _z(pnrel), _f(matom_left, np.complex128),
                          _f(matom_right, np.complex128), float(r_sph), ncrit, ngaussr,
                          _f(gr, np.float64), _f(gw, np.float64), _f(wy, np.complex128),
                          _f(cgnt, np.float64), cgnt.shape[0] - 1, pzz, pzj)
    return pzz, complex(pzj[0])
