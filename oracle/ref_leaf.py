"""ctypes bindings to oracle/_ref/liblsms_ref.so: the reference's own leaf
routines of the tau00 path (makegij_c.f, setgij.f, relmtrx.f, clebsch.f,
cgaunt_c.f90, gaunt.f, ifacts_c.f, congauss_c.f, spin_trafo.f,
associatedLegendreFunction.cpp, quadrature.cpp),
compiled from their sources by `make -C oracle ref`.

TEST INFRASTRUCTURE ONLY: used to pin the oracle's restatement against the
reference.  The library exists only where the reference tree was available
at build time; callers skip when `available()` is False.

Fortran character arguments carry a hidden trailing length (size_t).
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_ref", "liblsms_ref.so")
_lib = None

_zp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="F_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")
_ISTOP = C.create_string_buffer(b" " * 32, 33)


def available():
    return os.path.exists(_LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(_LIB_PATH)
    return _lib


def _i(v):
    return C.byref(C.c_int(int(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def _zarr(*vals):
    return np.asfortranarray(np.array(vals, dtype=np.complex128))


def lofk_mofk(lmax):
    lofk, mofk = [], []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            lofk.append(l)
            mofk.append(m)
    return (np.asfortranarray(np.array(lofk, dtype=np.int32)),
            np.asfortranarray(np.array(mofk, dtype=np.int32)))


def gauss_legendre_points(n, x1=-1.0, x2=1.0):
    x = np.zeros(n, order="F")
    w = np.zeros(n, order="F")
    lib().gauss_legendre_points_(_d(x1), _d(x2), x.ctypes.data_as(C.c_void_p),
                                 w.ctypes.data_as(C.c_void_p), _i(n))
    return x, w


def spin_trafo(evec):
    """Misc/spin_trafo.f: the spin rotation u and u^dagger (4 complex each,
    column-major 2 x 2) of a moment direction."""
    ev = np.asfortranarray(np.asarray(evec, dtype=np.float64))
    u = np.zeros(4, dtype=np.complex128, order="F")
    ud = np.zeros(4, dtype=np.complex128, order="F")
    lib().spin_trafo_(ev.ctypes.data_as(C.c_void_p), u.ctypes.data_as(C.c_void_p),
                      ud.ctypes.data_as(C.c_void_p))
    return u, ud


def plm_normalized(lmax, x):
    p = np.zeros((lmax + 1) * (lmax + 2) // 2, order="F")
    lib().plm_normalized_(_i(lmax), _d(x), p.ctypes.data_as(C.c_void_p))
    return p


def cgaunt(lmax):
    """Misc/cgaunt_c.f90 as called by GauntCoeficients::init (Coeficients.hpp:42-91)."""
    kkr = (lmax + 1) ** 2
    nd = (2 * lmax + 1) * (2 * lmax + 2) // 2
    clm = np.ones(nd, order="F")
    plmg = np.zeros((nd, 2 * lmax + 1), order="F")
    tg = np.zeros(2 * (2 * lmax + 1), order="F")
    wg = np.zeros(2 * (2 * lmax + 1), order="F")
    cg = np.zeros((lmax + 1, kkr, kkr), order="F")
    lofk, mofk = lofk_mofk(2 * lmax)
    iprint = C.c_int(-1)
    f = lib().cgaunt_
    f(_i(lmax), clm.ctypes.data_as(C.c_void_p), plmg.ctypes.data_as(C.c_void_p),
      tg.ctypes.data_as(C.c_void_p), wg.ctypes.data_as(C.c_void_p),
      cg.ctypes.data_as(C.c_void_p), lofk.ctypes.data_as(C.c_void_p),
      mofk.ctypes.data_as(C.c_void_p), C.byref(iprint), _ISTOP, C.c_size_t(32))
    return cg


def ifacts(lmax):
    """Misc/ifacts_c.f."""
    kkr = (lmax + 1) ** 2
    illp = np.zeros((kkr, kkr), dtype=np.complex128, order="F")
    ilp1 = np.zeros(2 * lmax + 1, dtype=np.complex128, order="F")
    iprint = C.c_int(-1)
    lib().ifacts_(_i(lmax), illp.ctypes.data_as(C.c_void_p), ilp1.ctypes.data_as(C.c_void_p),
                  C.byref(iprint), _ISTOP, C.c_size_t(32))
    return illp, ilp1


def makegij(lmaxi, lmaxj, lmax, prel, rij, cgnt=None, illp=None, ilp1=None):
    """MultipleScattering/makegij_c.f, with the tables of the reference."""
    if cgnt is None:
        cgnt = cgaunt(lmax)
    if illp is None:
        illp, ilp1 = ifacts(lmax)
    kkri, kkrj = (lmaxi + 1) ** 2, (lmaxj + 1) ** 2
    kkrsz = (lmax + 1) ** 2
    ndlj = (2 * lmax + 1) ** 2
    ndlm = (2 * lmax + 1) * (2 * lmax + 2) // 2
    lofk, mofk = lofk_mofk(2 * lmax)
    lend = lmaxi + lmaxj
    sinmp = np.zeros(lend + 1, order="F")
    cosmp = np.zeros(lend + 1, order="F")
    clm = np.ones(ndlm, order="F")
    plm = np.zeros(ndlm, order="F")
    hfn = np.zeros(lend + 2, dtype=np.complex128, order="F")
    dlm = np.zeros(ndlj, dtype=np.complex128, order="F")
    gij = np.zeros((kkri, kkrj), dtype=np.complex128, order="F")
    r = np.asfortranarray(np.asarray(rij, dtype=np.float64))
    pz = _zarr(prel)
    pi4 = C.c_double(4.0 * 2.0 * np.arcsin(1.0))
    iprint = C.c_int(-1)
    cg = np.asfortranarray(cgnt)
    il = np.asfortranarray(illp)
    i1 = np.asfortranarray(ilp1)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib().makegij_(_i(lmaxi), _i(kkri), _i(lmaxj), _i(kkrj), _i(lmax), _i(kkrsz), _i(ndlj),
                   _i(ndlm), vp(pz), vp(r), vp(sinmp), vp(cosmp), vp(clm), vp(plm), vp(cg),
                   _i(lmax), vp(lofk), vp(mofk), vp(i1), vp(il), vp(hfn), vp(dlm), vp(gij),
                   C.byref(pi4), C.byref(iprint), _ISTOP, C.c_size_t(32))
    return gij


def setgij(gij, n_spin_cant, energy):
    """MultipleScattering/setgij.f, non-relativistic branch."""
    kkr1, kkr2 = gij.shape
    b = np.zeros((kkr1 * n_spin_cant, kkr2 * n_spin_cant), dtype=np.complex128, order="F")
    g = np.asfortranarray(gij, dtype=np.complex128)
    psq = _zarr(0.0)
    ce = _zarr(energy)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib().setgij_(vp(g), vp(b), _i(kkr1), _i(kkr1 * n_spin_cant), _i(kkr2),
                  _i(kkr2 * n_spin_cant), _i(n_spin_cant), _i(0), vp(psq), vp(ce))
    return b


def clebsch():
    """Misc/clebsch.f: fills the common block /cgc/ (Misc/cgc.h); returns
    copies of u1, u2, ind1, ind2 (500 each)."""
    lib().clebsch_()

    class _Cgc(C.Structure):
        _fields_ = [("u1", C.c_double * 500), ("u2", C.c_double * 500),
                    ("ind1", C.c_int * 500), ("ind2", C.c_int * 500)]
    c = _Cgc.in_dll(lib(), "cgc_")
    return (np.array(c.u1[:]), np.array(c.u2[:]), np.array(c.ind1[:], dtype=np.int32),
            np.array(c.ind2[:], dtype=np.int32))


def relmtrx(a):
    """MultipleScattering/relmtrx.f (after clebsch_ has filled /cgc/)."""
    lib().clebsch_()
    kkr1, kkr2 = a.shape
    g = np.asfortranarray(a, dtype=np.complex128)
    b = np.zeros((2 * kkr1, 2 * kkr2), dtype=np.complex128, order="F")
    lib().relmtrx_(g.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p), _i(kkr1),
                   _i(kkr2))
    return b


def setgij_rel(gij, energy, prel):
    """MultipleScattering/setgij.f, relativistic branch (nrel_rel = 1):
    (prel^2 / E) relmtrx(gij)."""
    lib().clebsch_()
    kkr = gij.shape[0]
    b = np.zeros((2 * kkr, 2 * kkr), dtype=np.complex128, order="F")
    g = np.asfortranarray(gij, dtype=np.complex128)
    psq = _zarr(prel * prel)
    ce = _zarr(energy)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib().setgij_(vp(g), vp(b), _i(kkr), _i(2 * kkr), _i(kkr), _i(2 * kkr), _i(2), _i(1),
                  vp(psq), vp(ce))
    return b


def congauss(ebot, etop, eibot, npts):
    """Misc/congauss_c.f as called by buildEnergyContour (energyContourIntegration.cpp:106-112)."""
    ip = npts + 2
    egrd = np.zeros(ip, dtype=np.complex128, order="F")
    dele = np.zeros(ip, dtype=np.complex128, order="F")
    nume = C.c_int(0)
    pi = C.c_double(2.0 * np.arcsin(1.0))
    iprint = C.c_int(-1)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib().congauss_(_d(ebot), _d(etop), _d(eibot), vp(egrd), vp(dele), _i(npts), C.byref(nume),
                    C.byref(pi), _i(ip), C.byref(iprint), _ISTOP, C.c_size_t(32))
    return egrd[: nume.value], dele[: nume.value]


_SS_PATH = os.path.join(_HERE, "_ref", "libss_ref.so")
_ss = None


def ss_available():
    return os.path.exists(_SS_PATH)


def single_scatterer_nonrel(nrelv, lmax, energy, prel_, vr, r, jmt, jws, r_sph, clight=274.072):
    """SingleSite/single_scatterer_nonrel.f as called by
    calculateSingleScattererSolution (SingleSite/SingleSiteScattering.cpp:17-27).
    Returns (matom[lmax+1], tmat[kkrsz,kkrsz], zlr, jlr)."""
    global _ss
    if _ss is None:
        _ss = C.CDLL(_SS_PATH)
    vr = np.asfortranarray(np.asarray(vr, dtype=np.float64))
    r = np.asfortranarray(np.asarray(r, dtype=np.float64))
    iprpts = len(r)
    kkrsz = (lmax + 1) ** 2
    tmat = np.zeros((kkrsz, kkrsz), dtype=np.complex128, order="F")
    matom = np.zeros(lmax + 1, dtype=np.complex128, order="F")
    zlr = np.zeros((iprpts, lmax + 1), dtype=np.complex128, order="F")
    jlr = np.zeros((iprpts, lmax + 1), dtype=np.complex128, order="F")
    e = _zarr(energy)
    pr = _zarr(prel_)
    pn = _zarr(prel_)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    h = np.log(r[1] / r[0])
    _ss.single_scatterer_nonrel_(_i(nrelv), _d(clight), _i(lmax), _i(kkrsz), vp(e), vp(pr),
                                 vp(pn), vp(vr), vp(r), _d(h), _i(jmt), _i(jws), vp(tmat),
                                 vp(matom), vp(zlr), vp(jlr), _d(r_sph), _i(iprpts), _i(-1),
                                 _ISTOP, C.c_size_t(32))
    return matom, tmat, zlr, jlr


def _ssl():
    global _ss
    if _ss is None:
        _ss = C.CDLL(_SS_PATH)
    return _ss


def cnewint(x, f, ip0):
    """Misc/cnewint.f: running integral of f x^ip0 on the mesh x."""
    x = np.asfortranarray(np.asarray(x, dtype=np.float64))
    f = np.asfortranarray(np.asarray(f, dtype=np.complex128))
    g = np.zeros(len(x), dtype=np.complex128, order="F")
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    _ssl().cnewint_(_i(len(x)), vp(x), vp(f), vp(g), _i(ip0))
    return g


def cinterp(x, f, xs):
    """Misc/cinterp.f: value of f at xs."""
    x = np.asfortranarray(np.asarray(x, dtype=np.float64))
    f = np.asfortranarray(np.asarray(f, dtype=np.complex128))
    work = np.zeros(len(x), dtype=np.float64, order="F")
    ps = np.zeros(1, dtype=np.complex128)
    dps = np.zeros(1, dtype=np.complex128)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    _ssl().cinterp_(vp(x), vp(f), _i(len(x)), _d(xs), vp(ps), vp(dps), vp(work))
    return ps[0]


def mdosms(zj_flag, zz, zj, tau):
    """MultipleScattering/mdosms_c.f: -(sum zz tau - zj)/pi."""
    zz = np.asfortranarray(zz, dtype=np.complex128)
    tau = np.asfortranarray(tau, dtype=np.complex128)
    zjv = _zarr(zj)
    dos = np.zeros(1, dtype=np.complex128)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    _ssl().mdosms_(_i(zj_flag), _i(zz.size), vp(dos), vp(zz), vp(zjv), vp(tau),
                   _d(np.pi), _i(-1), _ISTOP, C.c_size_t(32))
    return dos[0]


def mgreen(zj_flag, lmax, jws, zlr, jlr, tau):
    """MultipleScattering/mgreen_c.f: green(r) for r < jws (left = right)."""
    zlr = np.asfortranarray(zlr, dtype=np.complex128)
    jlr = np.asfortranarray(jlr, dtype=np.complex128)
    tau = np.asfortranarray(tau, dtype=np.complex128)
    green = np.zeros(jws, dtype=np.complex128, order="F")
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    _ssl().mgreen_(_i(zj_flag), _i(tau.shape[0]), _i(lmax), _i(jws), _i(zlr.shape[0]), vp(zlr),
                   vp(zlr), vp(jlr), vp(green), vp(tau), _i(-1), _ISTOP, C.c_size_t(32))
    return green


def inter_m(zj_flag, lmax, pnrel, matom_left, matom_right, r_sph, cell, cgnt):
    """VORPOL/inter.f (inter_m -> interstitial): (pzz, pzj, pzj_full)."""
    ncrit, ngaussr, gr, gw, wy = cell
    kkr = (lmax + 1) ** 2
    lcg = cgnt.shape[0] - 1
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    ml = np.asfortranarray(matom_left, dtype=np.complex128)
    mr = np.asfortranarray(matom_right, dtype=np.complex128)
    g = np.asfortranarray(gr, dtype=np.float64)
    w = np.asfortranarray(gw, dtype=np.float64)
    y = np.asfortranarray(wy, dtype=np.complex128)
    cg = np.asfortranarray(cgnt, dtype=np.float64)
    pzz = np.zeros((kkr, kkr), dtype=np.complex128, order="F")
    pzj = np.zeros(1, dtype=np.complex128)
    pzf = np.zeros(kkr, dtype=np.complex128)
    pn = _zarr(pnrel)
    _ssl().inter_m_(_i(zj_flag), _i(lmax), vp(pn), vp(ml), vp(mr), _d(r_sph), _i(ngaussr),
                    vp(cg), _i(lcg), vp(pzz), vp(pzj), vp(pzf), _i(ncrit), vp(g), vp(w), vp(y),
                    _i(0), _ISTOP, C.c_size_t(32))
    return pzz, complex(pzj[0]), pzf
