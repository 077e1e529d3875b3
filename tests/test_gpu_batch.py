"""Uneven grouped batches (the 4096-site bcc alloy workload): schedules,
hipGraph replay and the autotuner, through the C ABI.

The schedule and the graph change which workgroup computes which tile and
how the launches reach the GPU, never the arithmetic of a solve, so results
must be bitwise identical across schedules, between graph replays and eager
launches, and between a large batch and the same solves in small batches.
Against LAPACK on the oracle's matrices the FP64 bar is 1e-10 relative.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from lsms_amd import capi
from lsms_amd import configs as CF
from tests import adjoint_ref as R

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def c4():
    s = CF.config_c4()
    ctx = capi.Context(0)
    ctx.load_system(s)
    return s, ctx


def pick(s, k):
    """k local atoms spanning the smallest to the largest KKR size."""
    n = np.array([s.nrmat_ns(i) for i in range(len(s.atoms))])
    order = np.argsort(n, kind="stable")
    return [int(order[j]) for j in np.linspace(0, len(order) - 1, k).round().astype(int)]


def lapack_tau00(s, a, ie):
    m = R.oracle_m(s, a, ie)
    rhs = np.zeros((m.shape[0], s.kkrsz_ns), dtype=np.complex128)
    rhs[:s.kkrsz_ns] = R.oracle_t0(s, a, ie)
    return sla.solve(m, rhs)[:s.kkrsz_ns]


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_MIXED_FP32, capi.PREC_MIXED_F16])
def test_schedules_bitwise_identical(c4, precision):
    s, ctx = c4
    atoms = pick(s, 6)
    iies = [3, 17, 31, 8, 25, 0]
    pr = [s.prel(i) for i in iies]
    outs = []
    for flags in range(8):
        ctx.set_schedule(flags)
        tau, info = ctx.tau00_batched(atoms, iies, pr, 16, precision=precision)
        assert np.all(info == 0)
        outs.append(tau)
    ctx.set_schedule(0)
    for t in outs[1:]:
        assert np.array_equal(t, outs[0])
    for b in (0, len(atoms) - 1):  # smallest and largest matrix
        ref = lapack_tau00(s, atoms[b], iies[b])
        assert rel(outs[0][b], ref) < 1e-10


def test_graph_replay_matches_eager(c4):
    s, ctx = c4
    atoms = pick(s, 8)
    rng = np.random.default_rng(7)
    ctx.set_schedule(capi.SCHED_RIGHT_ALIGN | capi.SCHED_TWO_STREAMS)
    for rep in range(3):
        iies = rng.integers(0, len(s.energies), len(atoms))
        pr = [s.prel(i) for i in iies]
        ctx.set_graph(False)
        eager, _ = ctx.tau00_batched(atoms, iies, pr, 16)
        ctx.set_graph(True)
        g1, info = ctx.tau00_batched(atoms, iies, pr, 16)
        # eager work between replays
        ctx.set_graph(False)
        ctx.tau00_batched(atoms[:2], iies[:2], pr[:2], 16)
        ctx.set_graph(True)
        g2, _ = ctx.tau00_batched(atoms, iies, pr, 16)
        assert np.all(info == 0)
        assert np.array_equal(g1, eager) and np.array_equal(g2, eager), rep
    ctx.set_graph(False)
    ctx.set_schedule(0)


def test_mixed_graph_replay_matches_eager(c4):
    """Mixed split-FP16 path, whose captured graph also holds the refinement
    sweeps: replays (with eager work between them) equal the eager result
    bitwise, and a change of the refinement limits is honoured by the graph
    (one sweep allowed: every solve falls back to the FP64 redo, as eager)."""
    s, ctx = c4
    atoms = pick(s, 8)
    iies = [1, 9, 17, 25, 4, 12, 20, 28]
    pr = [s.prel(i) for i in iies]
    mixed = capi.PREC_MIXED_F16
    ctx.set_schedule(capi.SCHED_RIGHT_ALIGN | capi.SCHED_TWO_STREAMS)
    ctx.set_graph(False)
    eager, info = ctx.tau00_batched(atoms, iies, pr, 16, precision=mixed)
    it_eager, _ = ctx.last_stats(len(atoms))
    assert np.all(info == 0) and np.all(it_eager >= 2)
    ctx.set_graph(True)
    g1, _ = ctx.tau00_batched(atoms, iies, pr, 16, precision=mixed)
    ctx.set_graph(False)
    ctx.tau00_batched(atoms[:3], iies[:3], pr[:3], 16, precision=mixed)
    ctx.set_graph(True)
    g2, _ = ctx.tau00_batched(atoms, iies, pr, 16, precision=mixed)
    it_graph, _ = ctx.last_stats(len(atoms))
    assert np.array_equal(g1, eager) and np.array_equal(g2, eager)
    assert np.array_equal(it_graph, it_eager)
    ctx.set_refinement(max_sweeps=1, tol=1e-12)
    g3, info3 = ctx.tau00_batched(atoms, iies, pr, 16, precision=mixed)
    it3, _ = ctx.last_stats(len(atoms))
    ctx.set_graph(False)
    e3, _ = ctx.tau00_batched(atoms, iies, pr, 16, precision=mixed)
    ctx.set_refinement()
    ctx.set_schedule(0)
    # one sweep: no solve converges, the FP64 redo cannot either, and every
    # solve is taken by the column-pivoted fallback (info 0, iters -2)
    assert np.all(info3 == 0) and np.all(it3 == -2)
    assert np.array_equal(g3, e3)
    for b in (0, len(atoms) - 1):
        assert rel(g3[b], lapack_tau00(s, atoms[b], iies[b])) < 1e-10
    for b in (0, len(atoms) - 1):
        assert rel(eager[b], lapack_tau00(s, atoms[b], iies[b])) < 1e-10


def test_autotune_then_same_results(c4):
    s, ctx = c4
    atoms = pick(s, 12)
    iies = list(range(12))
    pr = [s.prel(i) for i in iies]
    ctx.set_schedule(0)
    before, _ = ctx.tau00_batched(atoms, iies, pr, 16)
    best, ms = ctx.autotune(atoms, iies, pr)
    assert 0 <= best <= 7 and ctx.schedule() == best
    assert np.all(ms > 0)
    after, _ = ctx.tau00_batched(atoms, iies, pr, 16)
    assert np.array_equal(before, after)
    ctx.set_schedule(0)


@pytest.mark.parametrize("group", [4, 16, 8])
def test_split_group_sizes(c4, group):
    """Split-FP16 path with 4, 8 or 16 steps per trailing update (the fused
    panels of a group apply up to 15 earlier steps): uneven solves meet
    LAPACK at 1e-10, the
    schedules stay bitwise identical at each size, and a graph replay equals
    the eager result."""
    s, ctx = c4
    atoms = pick(s, 6)
    iies = [2, 19, 30, 7, 24, 13]
    pr = [s.prel(i) for i in iies]
    mixed = capi.PREC_MIXED_F16
    ctx.set_split_group(group)
    assert ctx.split_group()[0] == group
    outs = []
    for flags in (0, capi.SCHED_RIGHT_ALIGN | capi.SCHED_XCD_SWIZZLE):
        ctx.set_schedule(flags)
        tau, info = ctx.tau00_batched(atoms, iies, pr, 16, precision=mixed)
        assert np.all(info == 0)
        outs.append(tau)
    assert np.array_equal(outs[0], outs[1])
    ctx.set_graph(True)
    g1, _ = ctx.tau00_batched(atoms, iies, pr, 16, precision=mixed)
    g2, _ = ctx.tau00_batched(atoms, iies, pr, 16, precision=mixed)
    ctx.set_graph(False)
    assert np.array_equal(g1, outs[1]) and np.array_equal(g2, outs[1])
    ctx.set_schedule(0)
    ctx.set_split_group(8)
    for b in range(len(atoms)):
        err = rel(outs[0][b], lapack_tau00(s, atoms[b], iies[b]))
        assert err < 1e-10, (group, b, err)


def test_split_group_rejects_other_sizes(c4):
    s, ctx = c4
    for g in (0, 2, 6, 32):
        with pytest.raises(capi.LsmsxError):
            ctx.set_split_group(g)
    assert ctx.split_group()[0] == 8


def test_autotune_chooses_split_group(c4):
    """At the split-FP16 precision the autotuner also times the group sizes
    4, 8 and 16 under its best schedule and keeps the fastest."""
    s, ctx = c4
    atoms = pick(s, 8)
    iies = list(range(8))
    pr = [s.prel(i) for i in iies]
    best, ms = ctx.autotune(atoms, iies, pr, precision=capi.PREC_MIXED_F16)
    group, gms = ctx.split_group()
    assert 0 <= best <= 7 and np.all(ms > 0)
    assert group in (4, 8, 16) and np.all(gms > 0)
    assert gms[(4, 8, 16).index(group)] == gms.min() or (group == 8 and gms[1] <= 1.02 * gms.min())
    tau, info = ctx.tau00_batched(atoms, iies, pr, 16, precision=capi.PREC_MIXED_F16)
    assert np.all(info == 0)
    for b in (0, len(atoms) - 1):
        assert rel(tau[b], lapack_tau00(s, atoms[b], iies[b])) < 1e-10
    ctx.set_schedule(0)
    ctx.set_split_group(8)


def test_full_alloy_batch_graph(c4):
    """All 64 atoms x 32 energies (2048 uneven solves) in one graph launch
    sequence: a sample checked bitwise against small batches, and the
    extremes against LAPACK."""
    s, ctx = c4
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    pr = [s.prel(i) for i in iies]
    ctx.set_schedule(capi.SCHED_RIGHT_ALIGN | capi.SCHED_TWO_STREAMS)
    ctx.set_graph(True)
    tau, info = ctx.tau00_batched(atoms, iies, pr, 16)
    tau2, info2 = ctx.tau00_batched(atoms, iies, pr, 16)  # replay
    ctx.set_graph(False)
    ctx.set_schedule(0)
    assert np.all(info == 0) and np.all(info2 == 0)
    assert np.array_equal(tau, tau2)
    rng = np.random.default_rng(1)
    sample = rng.choice(len(atoms), 12, replace=False)
    small, _ = ctx.tau00_batched(atoms[sample], iies[sample], [pr[i] for i in sample], 16)
    assert np.array_equal(small, tau[sample])
    for a in pick(s, 2):
        b = a * ne + 5
        assert rel(tau[b], lapack_tau00(s, a, 5)) < 1e-10
