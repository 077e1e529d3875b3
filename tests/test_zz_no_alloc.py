"""lsmsx_solve_tau00_matrix_noalloc (device buffers in and out, scratch in a
static per-GPU device array) against lsmsx_solve_tau00_matrix on the same
matrices -- needs a GPU.

The two entry points run the same LU kernels on the same solver-layout copy,
so their tau00 must be bitwise equal; both are also checked against LAPACK to
the FP64 bar of 1e-10 relative.  Two calls in a row, the second smaller,
show that nothing of the first call's scratch leaks into the second.
"""
import numpy as np
import pytest

from lsms_amd import capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TAU_RTOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _problem(n, k, seed):
    rng = np.random.default_rng(seed)
    A = np.eye(n) + 0.3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
    t0 = rng.normal(size=(k, k)) + 1j * rng.normal(size=(k, k))
    return A, t0


class _Hip:
    """hipMalloc/hipMemcpy/hipFree of the HIP runtime the library itself is
    linked against (the test, not the entry point, owns the buffers)."""

    def __init__(self):
        import ctypes
        self.C = ctypes
        self.lib = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.C.c_void_p()
        assert self.lib.hipMalloc(self.C.byref(p), max(a.nbytes, 1)) == 0
        self.held.append(p.value)
        assert self.lib.hipMemcpy(p, a.ctypes.data_as(self.C.c_void_p), a.nbytes, 1) == 0
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.lib.hipMemcpy(out.ctypes.data_as(self.C.c_void_p), self.C.c_void_p(ptr),
                                  out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.held:
            self.lib.hipFree(self.C.c_void_p(p))
        self.held = []


def _noalloc(ctx, A, t0, ldm=None):
    n, k = A.shape[0], t0.shape[0]
    ldm = ldm or n
    # column-major with leading dimension ldm: element (r, c) at r + ldm c,
    # i.e. the C-order array of the (n x ldm) transpose
    mf = np.zeros((n, ldm), dtype=np.complex128)
    mf[:, :n] = A.T
    hip = _Hip()
    try:
        dm = hip.put(mf)
        dt = hip.put(np.ascontiguousarray(t0.T))
        dtau = hip.put(np.full((k, k), complex(np.nan, np.nan)))
        dinfo = hip.put(np.array([-7], dtype=np.int32))
        ctx.solve_tau00_matrix_noalloc(n, k, dm, ldm, dt, dtau, dinfo)
        tau = hip.get(dtau, np.zeros((k, k), dtype=np.complex128)).T.copy()
        info = int(hip.get(dinfo, np.zeros(1, dtype=np.int32))[0])
        # the caller's matrix is not modified
        assert np.array_equal(hip.get(dm, mf), mf)
    finally:
        hip.free()
    return tau, info


def test_noalloc_matches_allocating_entry_point_twice():
    ctx = capi.Context(0)
    try:
        for n, k, seed, ldm in ((500, 16, 3, 512), (240, 16, 5, None)):
            A, t0 = _problem(n, k, seed)
            tau_ref, info_ref = ctx.solve_tau00_matrix(A, t0)
            tau, info = _noalloc(ctx, A, t0, ldm)
            print(f"noalloc n={n} k={k}: info {info}, max |diff| "
                  f"{np.abs(tau - tau_ref).max():.3e}", flush=True)
            assert info == info_ref == 0
            assert np.array_equal(tau, tau_ref)
            to, _ = O.solve_tau00_zgetrf(A, t0)
            assert rel(tau, to) < TAU_RTOL
    finally:
        ctx.close()


def test_noalloc_k32_and_pivoting():
    """k = 32 (spin-canted central block) on a matrix with a zero diagonal,
    which needs row exchanges inside the tiles."""
    rng = np.random.default_rng(17)
    n, k = 300, 32
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A[np.arange(n), np.arange(n)] = 0.0
    t0 = np.eye(k, dtype=complex)
    ctx = capi.Context(0)
    try:
        tau_ref, info_ref = ctx.solve_tau00_matrix(A, t0)
        tau, info = _noalloc(ctx, A, t0)
    finally:
        ctx.close()
    assert info == info_ref == 0
    assert np.array_equal(tau, tau_ref)


def test_noalloc_rejects_sizes_beyond_the_limit():
    """The limit is checked on the host before any device work."""
    ctx = capi.Context(0)
    try:
        with pytest.raises(capi.LsmsxError, match="no-allocation limit"):
            ctx.solve_tau00_matrix_noalloc(1100, 16, 1, 1100, 1, 1, 1)
    finally:
        ctx.close()
