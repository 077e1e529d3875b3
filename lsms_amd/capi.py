"""Python binding of the C ABI in include/lsmsx.h (ctypes).

The native library lsms_amd/liblsmsx.so is the only compute path: if it is
missing this module raises instead of falling back to anything else.
"""
import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsmsx.so")

PREC_FP64 = 0
PREC_MIXED_FP32 = 1
PREC_MIXED_F16 = 2
PREC_FP64_REFINED = 3
(KIND_ASSEMBLE, KIND_DIAG, KIND_PANEL, KIND_UPDATE, KIND_FINISH, KIND_IR_RESIDUAL,
 KIND_IR_SOLVE, KIND_INVERSE, KIND_ADJOINT, KIND_CROSS) = range(10)

EXPORTED = [
    "lsmsx_set_split_group", "lsmsx_get_split_group", "lsmsx_set_shared_structure_constants",
    "lsmsx_last_error", "lsmsx_ctx_create", "lsmsx_ctx_destroy", "lsmsx_set_tables",
    "lsmsx_compute_tables", "lsmsx_upload_tmats", "lsmsx_set_liz", "lsmsx_build_kkr",
    "lsmsx_tau00", "lsmsx_tau00_batched", "lsmsx_solve_tau00_matrix",
    "lsmsx_set_workspace_limit", "lsmsx_set_profiling", "lsmsx_reset_profile",
    "lsmsx_get_profile", "lsmsx_set_refinement", "lsmsx_last_batch_stats",
    "lsmsx_upload_dtmats", "lsmsx_tau_full", "lsmsx_tau00_adjoint", "lsmsx_set_schedule",
    "lsmsx_get_schedule", "lsmsx_set_graph", "lsmsx_autotune", "lsmsx_set_pivot_guard",
    "lsmsx_single_site_nonrel", "lsmsx_green_mt", "lsmsx_set_single_site_resident",
    "lsmsx_green_mt_resident", "lsmsx_tmats_from_single_site", "lsmsx_green_mt_canted",
    "lsmsx_tmats_from_single_site_canted", "lsmsx_solve_tau00_matrix_noalloc",
    "lsmsx_set_residual_form", "lsmsx_upload_tmats_rel", "lsmsx_green_interstitial",
]
INFO_SWEEP_STALL = -2          # per-solve info: refinement wait bound hit (tau00 invalid)
INFO_REFINE_UNCONVERGED = -3   # per-solve info: FP64 refinement did not converge (plain result)
RESIDUAL_AUTO = 0
RESIDUAL_MATRIX = 1
SCHED_RIGHT_ALIGN = 1
SCHED_XCD_SWIZZLE = 2
SCHED_TWO_STREAMS = 4

_lib = None
_vp = C.c_void_p


def load():
    """Load liblsmsx.so; raises RuntimeError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"native library missing: {LIB_PATH}; build it with "
                           "`python -m lsms_amd.build` (no fallback path exists)")
    L = C.CDLL(LIB_PATH)
    L.lsmsx_last_error.restype = C.c_char_p
    L.lsmsx_ctx_create.argtypes = [C.c_int, C.POINTER(_vp)]
    L.lsmsx_ctx_destroy.argtypes = [_vp]
    L.lsmsx_set_tables.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
    L.lsmsx_compute_tables.argtypes = [C.c_int, _vp, _vp, _vp]
    L.lsmsx_upload_tmats.argtypes = [_vp, _vp, C.c_long, C.c_int, C.c_long, C.c_int]
    L.lsmsx_set_liz.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp]
    L.lsmsx_build_kkr.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp,
                                  C.c_int]
    L.lsmsx_tau00.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int,
                              _vp, _vp]
    L.lsmsx_tau00_batched.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, C.c_int, C.c_int,
                                      C.c_int, _vp, _vp]
    L.lsmsx_solve_tau00_matrix.argtypes = [_vp, C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, _vp]
    L.lsmsx_solve_tau00_matrix_noalloc.argtypes = [_vp, C.c_int, C.c_int, _vp, C.c_int, _vp,
                                                   _vp, _vp]
    L.lsmsx_set_workspace_limit.argtypes = [_vp, C.c_long]
    L.lsmsx_set_profiling.argtypes = [_vp, C.c_int]
    L.lsmsx_reset_profile.argtypes = [_vp]
    L.lsmsx_get_profile.argtypes = [_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_long)]
    L.lsmsx_set_refinement.argtypes = [_vp, C.c_int, C.c_double]
    L.lsmsx_set_residual_form.argtypes = [_vp, C.c_int]
    L.lsmsx_upload_tmats_rel.argtypes = [_vp, _vp, C.c_long, C.c_int, C.c_long, C.c_int, _vp]
    L.lsmsx_green_interstitial.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                           C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp,
                                           _vp, _vp, _vp, _vp, _vp]
    L.lsmsx_last_batch_stats.argtypes = [_vp, C.c_int, _vp, _vp]
    L.lsmsx_upload_dtmats.argtypes = [_vp, _vp, C.c_long, C.c_int, C.c_long, C.c_int]
    L.lsmsx_tau_full.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp,
                                 C.c_int, _vp]
    L.lsmsx_set_schedule.argtypes = [_vp, C.c_int]
    L.lsmsx_set_pivot_guard.argtypes = [_vp, C.c_double]
    L.lsmsx_get_schedule.argtypes = [_vp, _vp]
    L.lsmsx_set_graph.argtypes = [_vp, C.c_int]
    L.lsmsx_autotune.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                 _vp, _vp]
    L.lsmsx_tau00_adjoint.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int,
                                      _vp, _vp, _vp, _vp, _vp]
    L.lsmsx_set_single_site_resident.argtypes = [_vp, C.c_int]
    L.lsmsx_tmats_from_single_site_canted.argtypes = [_vp, C.c_long, C.c_int, C.c_long, C.c_int,
                                                      C.c_int, _vp, _vp, _vp, _vp, _vp]
    L.lsmsx_green_mt_canted.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                        _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.lsmsx_tmats_from_single_site.argtypes = [_vp, C.c_long, C.c_int, C.c_long, C.c_int, _vp, _vp]
    L.lsmsx_green_mt_resident.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp,
                                          _vp, _vp, _vp]
    L.lsmsx_green_mt.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp,
                                 _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.lsmsx_single_site_nonrel.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp,
                                           _vp, _vp, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]
    _lib = L
    return L


class LsmsxError(RuntimeError):
    pass


def _ck(rc):
    if rc != 0:
        msg = load().lsmsx_last_error().decode()
        raise LsmsxError(f"lsmsx error {rc}: {msg}")


def _p(a):
    return a.ctypes.data_as(_vp)


def _lmax_of_tau(tau00, nsolve):
    """lmax of a (nsolve, kkrsz, kkrsz) stack of tau00 blocks, kkrsz = (lmax+1)^2;
    ValueError for any other shape."""
    sh = np.shape(tau00)
    if len(sh) != 3 or sh[0] != nsolve or sh[1] != sh[2]:
        raise ValueError("tau00 must be (nsolve, kkrsz, kkrsz)")
    nl = int(round(np.sqrt(sh[1])))
    if nl * nl != sh[1] or not 1 <= nl <= 4:
        raise ValueError("tau00 blocks must be (lmax+1)^2 square with lmax 0..3")
    return nl - 1


def compute_tables(lmax):
    """Gaunt numbers cgnt (lmax+1, kkr, kkr), illp (kkr, kkr), ilp1 (2lmax+1)."""
    kkr = (lmax + 1) ** 2
    cg = np.zeros((lmax + 1, kkr, kkr), order="F")
    illp = np.zeros((kkr, kkr), dtype=np.complex128, order="F")
    ilp1 = np.zeros(2 * lmax + 1, dtype=np.complex128)
    _ck(load().lsmsx_compute_tables(lmax, _p(cg), _p(illp), _p(ilp1)))
    return cg, illp, ilp1


class Context:
    """One device context: tables, t-matrix store and LIZ data stay resident."""

    def __init__(self, device=0):
        self._h = _vp()
        _ck(load().lsmsx_ctx_create(device, C.byref(self._h)))
        self.nsc = 1
        self.nspola = 1

    def close(self):
        if self._h:
            load().lsmsx_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tables(self, lmax, cgnt=None, illp=None, ilp1=None):
        if cgnt is None:
            cgnt, illp, ilp1 = compute_tables(lmax)
        cg = np.asfortranarray(cgnt, dtype=np.float64)
        il = np.asfortranarray(illp, dtype=np.complex128)
        i1 = np.ascontiguousarray(ilp1, dtype=np.complex128)
        _ck(load().lsmsx_set_tables(self._h, lmax, _p(cg), _p(il), _p(i1)))

    def green_interstitial(self, lmax, prels, matom_left, matom_right, tau00, cells, cell_of,
                           r_sph, zj_flag=1, n_spin_pola=1):
        """Interstitial cell integrals and DOS: (dos (n,), pzz (n, kk, kk) row-major
        blocks, pzj (n,)).  cells: list of (ncrit, ngaussr, grwylm (ngaussr, nreg),
        gwwylm, wylm (nw, ngaussr, nreg)) sharing ngaussr; cell_of (n,); r_sph per cell."""
        n = len(prels)
        kk = (lmax + 1) ** 2
        nw = (2 * lmax + 1) * (lmax + 1)
        ngaussr = cells[0][1]
        maxreg = max(cl[0] - 1 for cl in cells)
        nc = len(cells)
        gr = np.zeros((nc, maxreg, ngaussr))
        gw = np.zeros((nc, maxreg, ngaussr))
        wy = np.zeros((nc, maxreg, ngaussr, nw), dtype=np.complex128)
        for k, (ncrit, ng, g, w, y) in enumerate(cells):
            assert ng == ngaussr
            gr[k, :ncrit - 1] = np.asarray(g).T
            gw[k, :ncrit - 1] = np.asarray(w).T
            wy[k, :ncrit - 1] = np.transpose(np.asarray(y), (2, 1, 0))
        ncrit = np.array([cl[0] for cl in cells], dtype=np.int32)
        pr = np.ascontiguousarray(prels, dtype=np.complex128)
        ml = np.ascontiguousarray(matom_left, dtype=np.complex128).reshape(n, lmax + 1)
        mr = np.ascontiguousarray(matom_right, dtype=np.complex128).reshape(n, lmax + 1)
        tau = np.ascontiguousarray(np.transpose(np.asarray(tau00, dtype=np.complex128), (0, 2, 1)))
        co = np.ascontiguousarray(cell_of, dtype=np.int32)
        rs = np.ascontiguousarray(r_sph, dtype=np.float64)
        dos = np.zeros(n, dtype=np.complex128)
        pzz = np.zeros((n, kk, kk), dtype=np.complex128)
        pzj = np.zeros(n, dtype=np.complex128)
        _ck(load().lsmsx_green_interstitial(self._h, n, lmax, zj_flag, n_spin_pola, _p(pr), _p(ml),
                                            _p(mr), nc, _p(co), _p(rs), _p(ncrit), ngaussr,
                                            maxreg, _p(np.ascontiguousarray(gr)),
                                            _p(np.ascontiguousarray(gw)),
                                            _p(np.ascontiguousarray(wy)), _p(tau), _p(pzz),
                                            _p(pzj), _p(dos)))
        return dos, np.ascontiguousarray(np.transpose(pzz, (0, 2, 1))), pzj

    def upload_tmats_rel(self, tmat_store, blk, group, energies):
        """Relativistic (kappa-mu) t-matrix store of `group` energy slots
        with their complex energies."""
        ts = np.asfortranarray(tmat_store, dtype=np.complex128)
        en = np.ascontiguousarray(energies, dtype=np.complex128)
        assert len(en) == group
        _ck(load().lsmsx_upload_tmats_rel(self._h, _p(ts), ts.shape[0], ts.shape[1], blk, group,
                                          _p(en)))

    def upload_tmats(self, tmat_store, blk, group):
        ts = np.asfortranarray(tmat_store, dtype=np.complex128)
        _ck(load().lsmsx_upload_tmats(self._h, _p(ts), ts.shape[0], ts.shape[1], blk, group))

    def set_liz(self, atom_id, liz_pos, liz_lmax, liz_store):
        pos = np.asfortranarray(liz_pos, dtype=np.float64)
        lm = np.ascontiguousarray(liz_lmax, dtype=np.int32)
        st = np.ascontiguousarray(liz_store, dtype=np.int32)
        _ck(load().lsmsx_set_liz(self._h, atom_id, pos.shape[1], _p(pos), _p(lm), _p(st)))

    def load_system(self, system):
        """Upload tables, t-matrix store and every atom's LIZ of a configs.System."""
        self.set_tables(system.maxlmax)
        self.upload_tmats(system.tmatStore, system.blkSizeTmatStore, len(system.energies))
        for a, at in enumerate(system.atoms):
            self.set_liz(a, at.LIZPos, at.LIZlmax, at.LIZStoreIdx)
        self.nsc = system.n_spin_cant
        self.nspola = system.n_spin_pola

    def build_kkr(self, atom_id, iie, prel, nrmat_ns, ispin=0):
        m = np.zeros((nrmat_ns, nrmat_ns), dtype=np.complex128, order="F")
        pr = np.array([complex(prel)], dtype=np.complex128)
        _ck(load().lsmsx_build_kkr(self._h, atom_id, ispin, iie, _p(pr), self.nsc, self.nspola,
                                   _p(m), nrmat_ns))
        return m

    def tau00(self, atom_id, iie, prel, kkrsz_ns, ispin=0, precision=PREC_FP64):
        tau = np.zeros((kkrsz_ns, kkrsz_ns), dtype=np.complex128, order="F")
        pr = np.array([complex(prel)], dtype=np.complex128)
        info = C.c_int(0)
        _ck(load().lsmsx_tau00(self._h, atom_id, ispin, iie, _p(pr), self.nsc, self.nspola,
                               precision, _p(tau), C.byref(info)))
        return tau, info.value

    def tau00_batched(self, atom_ids, iies, prels, kkrsz_ns, ispins=None, precision=PREC_FP64):
        n = len(atom_ids)
        a = np.ascontiguousarray(atom_ids, dtype=np.int32)
        e = np.ascontiguousarray(iies, dtype=np.int32)
        s = np.zeros(n, dtype=np.int32) if ispins is None else np.ascontiguousarray(ispins, dtype=np.int32)
        pr = np.ascontiguousarray(prels, dtype=np.complex128)
        tau = np.zeros((n, kkrsz_ns, kkrsz_ns), dtype=np.complex128)
        info = np.zeros(n, dtype=np.int32)
        _ck(load().lsmsx_tau00_batched(self._h, n, _p(a), _p(s), _p(e), _p(pr), self.nsc,
                                       self.nspola, precision, _p(tau), _p(info)))
        # each block is column-major kkrsz_ns x kkrsz_ns
        return np.ascontiguousarray(np.transpose(tau, (0, 2, 1))), info

    def solve_tau00_matrix(self, m, t0):
        mm = np.asfortranarray(m, dtype=np.complex128)
        tt = np.asfortranarray(t0, dtype=np.complex128)
        n, k = mm.shape[0], tt.shape[0]
        tau = np.zeros((k, k), dtype=np.complex128, order="F")
        info = C.c_int(0)
        _ck(load().lsmsx_solve_tau00_matrix(self._h, n, k, _p(mm), n, _p(tt), _p(tau),
                                            C.byref(info)))
        return tau, info.value

    def solve_tau00_matrix_noalloc(self, n, k, d_m, ldm, d_t0, d_tau00, d_info):
        """lsmsx_solve_tau00_matrix_noalloc on device addresses (integers):
        d_m n x n (leading dimension ldm, central block first), d_t0 k x k,
        d_tau00 k x k (written) and d_info one int32 (written), all
        column-major complex128.  Nothing is allocated on the device."""
        _ck(load().lsmsx_solve_tau00_matrix_noalloc(self._h, n, k, C.c_void_p(d_m), ldm,
                                                    C.c_void_p(d_t0), C.c_void_p(d_tau00),
                                                    C.c_void_p(d_info)))

    def upload_dtmats(self, dtmat_store, blk, group):
        ts = np.asfortranarray(dtmat_store, dtype=np.complex128)
        _ck(load().lsmsx_upload_dtmats(self._h, _p(ts), ts.shape[0], ts.shape[1], blk, group))

    def tau_full(self, atom_id, iie, prel, nrmat_ns, ispin=0):
        """tau = M^-1 T (nrmat_ns x nrmat_ns, reference site order) and info."""
        tau = np.zeros((nrmat_ns, nrmat_ns), dtype=np.complex128, order="F")
        pr = np.array([complex(prel)], dtype=np.complex128)
        info = C.c_int(0)
        _ck(load().lsmsx_tau_full(self._h, atom_id, ispin, iie, _p(pr), self.nsc, self.nspola,
                                  _p(tau), nrmat_ns, C.byref(info)))
        return tau, info.value

    def single_site_nonrel(self, nrelv, lmax, vr, r_mesh, jmt, jws, r_sph, energies, prels,
                           radial=True):
        """Single-site solve of npot potentials at ne energies on the GPU
        (single_scatterer_nonrel.f via semrel_c.f, solver 2).  vr, r_mesh:
        (npot, iprpts); jmt, jws: 1-based per potential; r_sph per potential.
        Returns (matom[ne, npot, lmax+1], tmat[ne, npot, kkrsz, kkrsz],
        zlr, jlr [ne, npot, lmax+1, iprpts] or None, status) where status 1
        means an integration step underflowed somewhere."""
        vr = np.ascontiguousarray(np.atleast_2d(vr), dtype=np.float64)
        rm = np.ascontiguousarray(np.atleast_2d(r_mesh), dtype=np.float64)
        npot, iprpts = vr.shape
        if rm.shape != vr.shape:
            raise ValueError("vr and r_mesh shapes differ")
        jm = np.ascontiguousarray(np.broadcast_to(jmt, (npot,)), dtype=np.int32)
        jw = np.ascontiguousarray(np.broadcast_to(jws, (npot,)), dtype=np.int32)
        rs = np.ascontiguousarray(np.broadcast_to(r_sph, (npot,)), dtype=np.float64)
        e = np.ascontiguousarray(np.atleast_1d(energies), dtype=np.complex128)
        pr = np.ascontiguousarray(np.atleast_1d(prels), dtype=np.complex128)
        ne, nl = len(e), lmax + 1
        if len(pr) != ne:
            raise ValueError("energies and prels lengths differ")
        kk = nl * nl
        matom = np.zeros((ne, npot, nl), dtype=np.complex128)
        tmat = np.zeros((ne, npot, kk, kk), dtype=np.complex128)
        zlr = np.zeros((ne, npot, nl, iprpts), dtype=np.complex128) if radial else None
        jlr = np.zeros((ne, npot, nl, iprpts), dtype=np.complex128) if radial else None
        rc = load().lsmsx_single_site_nonrel(
            self._h, nrelv, lmax, npot, iprpts, _p(vr), _p(rm), _p(jm), _p(jw), _p(rs), ne,
            _p(e), _p(pr), _p(matom), _p(tmat), _p(zlr) if radial else None,
            _p(jlr) if radial else None)
        if rc < 0:
            _ck(rc)
        # tmat blocks are column-major kkrsz x kkrsz
        return matom, np.swapaxes(tmat, 2, 3), zlr, jlr, rc

    def green_mt(self, lmax, tau00, zlr, jlr, pot_of, r_mesh, jmt, jws, r_sph, n_spin_pola=1,
                 with_green=True):
        """Muffin-tin DOS and spherical Green function of nsolve solves of one
        collinear spin (green_function.f, sphere part).  tau00: (nsolve,
        kkrsz, kkrsz); zlr, jlr: (nsolve, lmax+1, iprpts); pot_of: (nsolve,);
        r_mesh: (npot, iprpts).  Returns (dos[nsolve], green[nsolve, iprpts])."""
        rm = np.ascontiguousarray(np.atleast_2d(r_mesh), dtype=np.float64)
        npot, iprpts = rm.shape
        z = np.ascontiguousarray(zlr, dtype=np.complex128)
        j = np.ascontiguousarray(jlr, dtype=np.complex128)
        nsolve = z.shape[0]
        if z.shape != (nsolve, lmax + 1, iprpts) or j.shape != z.shape:
            raise ValueError("zlr/jlr must be (nsolve, lmax+1, iprpts)")
        if _lmax_of_tau(tau00, nsolve) != lmax:
            raise ValueError("tau00 must be (nsolve, (lmax+1)^2, (lmax+1)^2)")
        # column-major kkrsz x kkrsz blocks
        t = np.ascontiguousarray(np.swapaxes(np.asarray(tau00, dtype=np.complex128), 1, 2))
        po = np.ascontiguousarray(pot_of, dtype=np.int32)
        if po.shape != (nsolve,):
            raise ValueError("pot_of must hold one potential per solve")
        jm = np.ascontiguousarray(np.broadcast_to(jmt, (npot,)), dtype=np.int32)
        jw = np.ascontiguousarray(np.broadcast_to(jws, (npot,)), dtype=np.int32)
        rs = np.ascontiguousarray(np.broadcast_to(r_sph, (npot,)), dtype=np.float64)
        dos = np.zeros(nsolve, dtype=np.complex128)
        green = np.zeros((nsolve, iprpts), dtype=np.complex128) if with_green else None
        _ck(load().lsmsx_green_mt(self._h, nsolve, lmax, iprpts, n_spin_pola, npot, _p(po), _p(rm),
                                  _p(jm), _p(jw), _p(rs), _p(t), _p(z), _p(j), _p(dos),
                                  _p(green) if with_green else None))
        return dos, green

    def tmats_from_single_site(self, ld, num_store, blk, group, store_of_pot, spin_of_pot=None):
        """Fill the device t-matrix store from the last single_site_nonrel
        batch (zeroed first; energy e -> slot e)."""
        so = np.ascontiguousarray(store_of_pot, dtype=np.int32)
        sp = np.ascontiguousarray(np.zeros_like(so) if spin_of_pot is None else spin_of_pot,
                                  dtype=np.int32)
        _ck(load().lsmsx_tmats_from_single_site(self._h, ld, num_store, blk, group, _p(so), _p(sp)))

    def green_mt_canted(self, lmax, tau00_l, zlr, jlr, wxyz, pot_of, r_mesh, jmt, jws, r_sph,
                        with_green=True):
        """Spin-canted muffin-tin DOS and Green function.  tau00_l: (nsolve, 4,
        kkrsz, kkrsz); zlr, jlr: (nsolve, 2, lmax+1, iprpts); wxyz: (nsolve, 3, 4).
        Returns (dos[nsolve, 4], green[nsolve, 4, iprpts]) with parts Tr, x, y, z."""
        rm = np.ascontiguousarray(np.atleast_2d(r_mesh), dtype=np.float64)
        npot, iprpts = rm.shape
        z = np.ascontiguousarray(zlr, dtype=np.complex128)
        j = np.ascontiguousarray(jlr, dtype=np.complex128)
        nsolve = z.shape[0]
        if z.shape != (nsolve, 2, lmax + 1, iprpts) or j.shape != z.shape:
            raise ValueError("zlr/jlr must be (nsolve, 2, lmax+1, iprpts)")
        kk = (lmax + 1) ** 2
        if np.shape(tau00_l) != (nsolve, 4, kk, kk):
            raise ValueError("tau00_l must be (nsolve, 4, (lmax+1)^2, (lmax+1)^2)")
        if np.shape(wxyz) != (nsolve, 3, 4):
            raise ValueError("wxyz must be (nsolve, 3, 4)")
        t = np.ascontiguousarray(np.swapaxes(np.asarray(tau00_l, dtype=np.complex128), 2, 3))
        w = np.ascontiguousarray(wxyz, dtype=np.complex128)
        po = np.ascontiguousarray(pot_of, dtype=np.int32)
        if po.shape != (nsolve,):
            raise ValueError("pot_of must hold one potential per solve")
        jm = np.ascontiguousarray(np.broadcast_to(jmt, (npot,)), dtype=np.int32)
        jw = np.ascontiguousarray(np.broadcast_to(jws, (npot,)), dtype=np.int32)
        rs = np.ascontiguousarray(np.broadcast_to(r_sph, (npot,)), dtype=np.float64)
        dos = np.zeros((nsolve, 4), dtype=np.complex128)
        green = np.zeros((nsolve, 4, iprpts), dtype=np.complex128) if with_green else None
        _ck(load().lsmsx_green_mt_canted(self._h, nsolve, lmax, iprpts, npot, _p(po), _p(rm),
                                         _p(jm), _p(jw), _p(rs), _p(t), _p(z), _p(j), _p(w),
                                         _p(dos), _p(green) if with_green else None))
        return dos, green

    def tmats_from_single_site_canted(self, ld, num_store, blk, group, store_of_atom, pot_up,
                                      pot_dn, ubr, ubrd):
        """Spin-canted store fill from the last single_site_nonrel batch:
        trltog(t_up, t_dn) per atom with its (ubr, ubrd) (natom x 4 complex)."""
        so = np.ascontiguousarray(store_of_atom, dtype=np.int32)
        pu = np.ascontiguousarray(pot_up, dtype=np.int32)
        pd = np.ascontiguousarray(pot_dn, dtype=np.int32)
        u = np.ascontiguousarray(ubr, dtype=np.complex128).reshape(len(so), 4)
        ud = np.ascontiguousarray(ubrd, dtype=np.complex128).reshape(len(so), 4)
        _ck(load().lsmsx_tmats_from_single_site_canted(self._h, ld, num_store, blk, group, len(so),
                                                       _p(so), _p(pu), _p(pd), _p(u), _p(ud)))

    def set_single_site_resident(self, on):
        """Keep the radial functions of later single_site_nonrel calls on the
        device for green_mt_resident."""
        _ck(load().lsmsx_set_single_site_resident(self._h, 1 if on else 0))

    def green_mt_resident(self, tau00, pair_of, iprpts, n_spin_pola=1, with_green=True,
                          jmt_of_pot=None):
        """green_mt on the radial functions left on the device by the last
        single_site_nonrel call (residency on); pair_of[b] = e*npot + p;
        jmt_of_pot overrides the end of the sphere integrals (jws for ASA)."""
        po = np.ascontiguousarray(pair_of, dtype=np.int32)
        nsolve = len(po)
        lmax = _lmax_of_tau(tau00, nsolve)
        t = np.ascontiguousarray(np.swapaxes(np.asarray(tau00, dtype=np.complex128), 1, 2))
        dos = np.zeros(nsolve, dtype=np.complex128)
        green = np.zeros((nsolve, iprpts), dtype=np.complex128) if with_green else None
        jm = None if jmt_of_pot is None else np.ascontiguousarray(jmt_of_pot, dtype=np.int32)
        _ck(load().lsmsx_green_mt_resident(self._h, nsolve, lmax, int(iprpts), n_spin_pola,
                                           _p(po), None if jm is None else _p(jm), _p(t), _p(dos),
                                           _p(green) if with_green else None))
        return dos, green

    def tau00_adjoint(self, atom_ids, iies, energies, kkrsz_ns, ispins=None, lloyd=True,
                      tau00_bar=None, store_shape=None):
        """Adjoint route: (tau00 (n, k, k), lloyd (n,) or None, t-store gradient or None, info).
        tau00_bar: (n, k, k) cotangents (row-major blocks, like the returned tau00);
        store_shape: shape of the t-matrix store (for the gradient)."""
        n = len(atom_ids)
        a = np.ascontiguousarray(atom_ids, dtype=np.int32)
        e = np.ascontiguousarray(iies, dtype=np.int32)
        s = np.zeros(n, dtype=np.int32) if ispins is None else np.ascontiguousarray(ispins, dtype=np.int32)
        en = np.ascontiguousarray(energies, dtype=np.complex128)
        from .configs import prel_of
        pr = np.ascontiguousarray(prel_of(en), dtype=np.complex128)
        tau = np.zeros((n, kkrsz_ns, kkrsz_ns), dtype=np.complex128)
        ll = np.zeros(n, dtype=np.complex128) if lloyd else None
        info = np.zeros(n, dtype=np.int32)
        vb = grad = None
        if tau00_bar is not None:
            vb = np.ascontiguousarray(np.transpose(np.asarray(tau00_bar, dtype=np.complex128),
                                                   (0, 2, 1)))
            grad = np.zeros(store_shape, dtype=np.complex128, order="F")
        _ck(load().lsmsx_tau00_adjoint(self._h, n, _p(a), _p(s), _p(e), _p(en), _p(pr), self.nsc,
                                       self.nspola, _p(tau), _p(ll) if lloyd else None,
                                       _p(vb) if vb is not None else None,
                                       _p(grad) if grad is not None else None, _p(info)))
        return np.ascontiguousarray(np.transpose(tau, (0, 2, 1))), ll, grad, info

    def set_pivot_guard(self, threshold):
        _ck(load().lsmsx_set_pivot_guard(self._h, float(threshold)))

    def set_schedule(self, flags):
        _ck(load().lsmsx_set_schedule(self._h, int(flags)))

    def schedule(self):
        f = C.c_int(0)
        _ck(load().lsmsx_get_schedule(self._h, C.byref(f)))
        return f.value

    def set_graph(self, on):
        _ck(load().lsmsx_set_graph(self._h, 1 if on else 0))

    def autotune(self, atom_ids, iies, prels, ispins=None, precision=PREC_FP64):
        """Time the batch under the eight schedules; returns (best flags, ms[8])."""
        n = len(atom_ids)
        a = np.ascontiguousarray(atom_ids, dtype=np.int32)
        e = np.ascontiguousarray(iies, dtype=np.int32)
        s = np.zeros(n, dtype=np.int32) if ispins is None else np.ascontiguousarray(ispins, dtype=np.int32)
        pr = np.ascontiguousarray(prels, dtype=np.complex128)
        best = C.c_int(0)
        ms = np.zeros(8)
        _ck(load().lsmsx_autotune(self._h, n, _p(a), _p(s), _p(e), _p(pr), self.nsc, self.nspola,
                                  precision, C.byref(best), _p(ms)))
        return best.value, ms

    def set_shared_structure_constants(self, on):
        """Share G_ij(E) among the solves of a batch with equal LIZ geometry,
        energy slot and prel (default on; results are bitwise the same)."""
        _ck(load().lsmsx_set_shared_structure_constants(self._h, 1 if on else 0))

    def set_split_group(self, group):
        """Steps per trailing update of the split-FP16 path: 4, 8 or 16."""
        _ck(load().lsmsx_set_split_group(self._h, int(group)))

    def split_group(self):
        """(current split group, wall ms of groups 4, 8, 16 from the last
        split-FP16 autotune, -1 where not measured)."""
        g = C.c_int(0)
        ms = np.zeros(3)
        _ck(load().lsmsx_get_split_group(self._h, C.byref(g), _p(ms)))
        return g.value, ms

    def set_residual_form(self, form):
        """RESIDUAL_AUTO (structure constants once per site pair where the
        batch allows it) or RESIDUAL_MATRIX (always the FP64 M)."""
        _ck(load().lsmsx_set_residual_form(self._h, form))

    def set_refinement(self, max_sweeps=12, tol=1e-12):
        _ck(load().lsmsx_set_refinement(self._h, int(max_sweeps), float(tol)))

    def last_stats(self, n):
        """(refinement sweeps per solve: 0 plain, > 0 refined, -1 mixed solve redone in
        FP64, -2 solved by the column-pivoted fallback; smallest |pivot| of the
        tile-pivoted factorization of each solve)."""
        it = np.zeros(n, dtype=np.int32)
        mp = np.zeros(n, dtype=np.float64)
        _ck(load().lsmsx_last_batch_stats(self._h, n, _p(it), _p(mp)))
        return it, mp

    def set_workspace_limit(self, nbytes):
        _ck(load().lsmsx_set_workspace_limit(self._h, int(nbytes)))

    def set_profiling(self, on):
        _ck(load().lsmsx_set_profiling(self._h, 1 if on else 0))

    def reset_profile(self):
        _ck(load().lsmsx_reset_profile(self._h))

    def profile(self, kind):
        ms = C.c_double(0)
        n = C.c_long(0)
        _ck(load().lsmsx_get_profile(self._h, kind, C.byref(ms), C.byref(n)))
        return ms.value, n.value
