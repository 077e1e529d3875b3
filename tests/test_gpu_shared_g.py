"""Structure constants shared within a batch -- needs a GPU.

Solves whose atoms have bitwise equal LIZ geometry and the same energy slot
share G_ij(E): the first of each group evaluates it once per site pair and
every solve builds its blocks from the stored values.  Those are the FP64
values each solve would compute itself, so the matrices, and with them every
result, must be bitwise the same with sharing on and off, on every path that
uses it (plain FP64, mixed with the structure-constant residual, mixed with
the FP64 matrix), under the graph and the two-stream schedule, and on the
uneven alloy batch where only some atoms share a geometry.  Against LAPACK
the bar stays 1e-10.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from lsms_amd import capi
from lsms_amd import configs as CF
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def lapack_tau00(s, a, ie):
    at = s.atoms[a]
    m = O.build_kkr(s.maxlmax, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore,
                    s.blkSizeTmatStore, ie, s.prel(ie))
    t0 = O.t0(s.maxlmax, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, ie)
    rhs = np.zeros((m.shape[0], t0.shape[0]), complex)
    rhs[:t0.shape[0]] = t0
    return sl.solve(m, rhs)[:t0.shape[0]]


@pytest.fixture(scope="module")
def c2pair():
    s = CF.config_c2(npts=4, natoms=2)
    assert np.array_equal(s.atoms[0].LIZPos, s.atoms[1].LIZPos)
    ctx = capi.Context(0)
    ctx.load_system(s)
    yield s, ctx
    ctx.close()


def run(ctx, atoms, iies, prels, precision, share):
    ctx.set_shared_structure_constants(share)
    tau, info = ctx.tau00_batched(atoms, iies, prels, 16, precision=precision)
    ctx.set_shared_structure_constants(True)
    return tau, info


@pytest.mark.parametrize("precision,form", [(capi.PREC_FP64, None),
                                            (capi.PREC_MIXED_F16, capi.RESIDUAL_AUTO),
                                            (capi.PREC_MIXED_FP32, capi.RESIDUAL_AUTO),
                                            (capi.PREC_MIXED_F16, capi.RESIDUAL_MATRIX)])
def test_shared_equals_per_solve_bitwise(c2pair, precision, form):
    s, ctx = c2pair
    atoms = np.array([0, 1, 0, 1, 1, 0])
    iies = np.array([0, 0, 2, 2, 3, 1])   # three pairs share, two solves alone
    prels = np.array([s.prel(i) for i in iies])
    if form is not None:
        ctx.set_residual_form(form)
    shared, info1 = run(ctx, atoms, iies, prels, precision, True)
    alone, info2 = run(ctx, atoms, iies, prels, precision, False)
    ctx.set_residual_form(capi.RESIDUAL_AUTO)
    assert np.all(info1 == 0) and np.all(info2 == 0)
    assert np.array_equal(shared, alone)
    for b in (0, 3):
        ref = lapack_tau00(s, atoms[b], iies[b])
        assert np.linalg.norm(shared[b] - ref) / np.linalg.norm(ref) < 1e-10


def test_shared_under_graph_and_two_streams(c2pair):
    s, ctx = c2pair
    atoms = np.array([0, 1, 1, 0, 0, 1, 0, 1])
    iies = np.array([1, 1, 3, 3, 0, 0, 2, 2])
    prels = np.array([s.prel(i) for i in iies])
    eager, _ = run(ctx, atoms, iies, prels, capi.PREC_MIXED_F16, False)
    ctx.set_schedule(capi.SCHED_RIGHT_ALIGN | capi.SCHED_XCD_SWIZZLE | capi.SCHED_TWO_STREAMS)
    ctx.set_graph(True)
    g1, _ = ctx.tau00_batched(atoms, iies, prels, 16, precision=capi.PREC_MIXED_F16)
    ctx.set_graph(False)
    ctx.tau00_batched(atoms[:3], iies[:3], prels[:3], 16)   # eager work between replays
    ctx.set_graph(True)
    g2, _ = ctx.tau00_batched(atoms, iies, prels, 16, precision=capi.PREC_MIXED_F16)
    ctx.set_graph(False)
    ctx.set_schedule(0)
    assert np.array_equal(g1, eager) and np.array_equal(g2, eager)


def test_alloy_batch_with_partly_shared_geometry():
    """The uneven alloy batch: atoms whose LIZs differ (vacancies) evaluate
    their own structure constants, those that match share; bitwise equal to
    per-solve evaluation either way."""
    s = CF.config_c4()
    ctx = capi.Context(0)
    ctx.load_system(s)
    geo = {}
    for a, at in enumerate(s.atoms[:24]):
        geo.setdefault((at.LIZPos.tobytes(), at.LIZlmax.tobytes()), []).append(a)
    atoms = np.repeat(np.arange(24), 2)
    iies = np.tile(np.array([5, 17]), 24)
    prels = np.array([s.prel(i) for i in iies])
    for prec in (capi.PREC_FP64, capi.PREC_MIXED_F16):
        shared, i1 = run(ctx, atoms, iies, prels, prec, True)
        alone, i2 = run(ctx, atoms, iies, prels, prec, False)
        assert np.all(i1 == 0) and np.all(i2 == 0)
        assert np.array_equal(shared, alone), prec
    ctx.close()
    print(f"{len(geo)} distinct LIZ geometries among 24 alloy atoms")


def test_shared_fallback_keeps_each_solve_in_its_slot(c2pair):
    """FP64-refined solves allowed one sweep cannot converge, so each is
    solved again by the column-pivoted fallback (info 0, iters -2): with
    sharing the chunk runs in group order, and every fallback must still land
    in its own slot -- each result meets LAPACK for its own (atom, energy),
    bitwise the same with sharing on and off."""
    s, ctx = c2pair
    atoms = np.array([0, 1, 1, 0, 0, 1])
    iies = np.array([3, 1, 3, 1, 0, 0])
    prels = np.array([s.prel(i) for i in iies])
    ctx.set_refinement(max_sweeps=1, tol=1e-12)
    shared, info = run(ctx, atoms, iies, prels, capi.PREC_FP64_REFINED, True)
    iters, _ = ctx.last_stats(len(atoms))
    alone, _ = run(ctx, atoms, iies, prels, capi.PREC_FP64_REFINED, False)
    ctx.set_refinement()
    assert np.all(info == 0) and np.all(iters == -2)
    assert np.array_equal(shared, alone)
    for b in range(len(atoms)):
        ref = lapack_tau00(s, atoms[b], iies[b])
        assert np.linalg.norm(shared[b] - ref) / np.linalg.norm(ref) < 1e-10
