"""Throughput of the tau00 path on one MI355X.

One step = every (local atom, contour energy) tau00 solve of the workload:
KKR assembly M = 1 - t G(E) on the device plus the central-block inversion
tau00 = [M^-1 [t0;0]]_00.  Workload: fcc a = 6.82 bohr, LIZ of
135 sites (N = 2160), lmax = 3, 32 local atoms x 31 semicircle energies
(992 solves per step).  t-matrices, LIZ geometry and tables are resident on
the device before timing; only the 16x16 tau00 blocks come back.  For each
path (plain FP64; mixed with FP32 trailing updates; mixed with split-FP16
trailing updates) the library's autotuner first picks the batch schedule
(untimed), then the timed steps replay the captured hipGraph; the headline
is the fastest mixed path whose tau00 agree with the FP64 path to 1e-10.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tau00 solves/sec at lmax=3 LIZ=135: mixed+batched vs FP64, <=1e-10 rel err"
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (matrix) peak
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 (matrix) peak
FP16_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 (matrix) peak
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E


def solve_flops(n, k):
    """Algorithmic work of one tau00 solve: complex LU + two triangular solves
    with k right-hand sides, F = 8/3 N^3 + 8 N^2 k real flops."""
    return (8.0 / 3.0) * n ** 3 + 8.0 * n * n * k


def _groups(nt, group):
    """(steps in the group, trailing tiles per dimension) of every trailing
    update of one solve: groups of `group` elimination steps g .. g+ns-1 over
    steps 0 .. nt-2 (step nt-1 is the last tile), each followed by one update
    of the (nt-1-last)^2 tiles beyond the group."""
    for g in range(0, nt - 1, group):
        ns = min(group, nt - 1 - g)
        r = nt - 1 - (g + ns - 1)
        if r > 0:
            yield ns, r


def update_kernel_bytes(nt_list, c_bytes, panel_bytes, group):
    """Algorithmic HBM bytes of all trailing-update launches of one batch: each
    update reads and writes its r^2 C tiles once and reads the ns L column and
    ns U row panels of its group once (bytes per complex element: c_bytes for
    C, panel_bytes for the panel operands)."""
    total = 0.0
    for nt in nt_list:
        for ns, r in _groups(nt, group):
            total += r * r * 2 * 64 * 64 * c_bytes + 2 * ns * r * 64 * 64 * panel_bytes
    return total


def update_kernel_flops(nt_list, group):
    """Algorithmic flops of all trailing-update launches of one batch: the ns
    rank-64 panels of a group applied to the r^2 tiles beyond it, ns * 8*64^3
    flops per tile (8 flops per complex FMA)."""
    total = 0.0
    for nt in nt_list:
        for ns, r in _groups(nt, group):
            total += r * r * ns * 8.0 * 64 ** 3
    return total


PROFILES = [os.path.join(ROOT, "profiles", f) for f in
            ("k_lu_trail_fcc135_N2160.md", "k_lu_trail_h_fcc135_N2160.md")]


def profiled_value(field, kernel):
    """A per-launch figure of `kernel` on this workload recorded in the profile
    summaries as a `field[kernel]: value` line: `traffic_bytes_per_launch`
    (HBM bytes from the TCC_EA0_RDREQ request counts and WRITE_SIZE, checked
    against a copy kernel of known size), `mfma_util_counter` (MfmaUtil =
    SQ_VALU_MFMA_BUSY_CYCLES over GRBM_GUI_ACTIVE per XCD x 1024 SIMDs) and
    `mfma_flops_executed_per_launch` (512 x SQ_INSTS_VALU_MFMA_MOPS_<type>).
    The summaries hold the rocprofv3 commands these come from."""
    key = "%s[%s]:" % (field, kernel)
    for path in PROFILES:
        try:
            for line in open(path):
                if line.startswith(key):
                    return float(line.split(":")[1])
        except OSError:
            pass
    return None


def host_cpu():
    """(logical CPUs this process may use, CPU model) of the host."""
    try:
        ncpu = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        ncpu = os.cpu_count() or 1
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return ncpu, model


def cpu_baseline(system, atoms, iies, budget_s=(6.0, 8.0, 4.0)):
    """Host-core baseline of the reference's CPU model and LAPACK references.

    Each solve is the oracle's C assembly of M (calculateTauMatrix.cpp:57-315)
    plus LAPACK zgetrf/zgetrs (scipy's OpenBLAS; linearSolvers_CPU.cpp:98-123),
    on bounded samples of the workload's solves, timed three ways:
      (i)   one solve per thread over all host CPUs, sequential LAPACK: the
            reference's own model, an OpenMP loop over independent (atom,
            energy) solves (calculateTauMatrix.cpp:891-900);
      (ii)  one solve at a time, LAPACK threaded over all host CPUs;
      1-core one solve at a time on one thread.
    Returns (summary dict, {solve index: LAPACK tau00}); every solve timed in
    any mode is kept as a LAPACK reference for the accuracy gate."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import scipy.linalg as sl
    from threadpoolctl import threadpool_limits
    from oracle import oracle as O
    ncpu, model = host_cpu()
    refs = {}
    lock = threading.Lock()
    order = [(37 * q) % len(atoms) for q in range(len(atoms))]

    def one(idx):
        a, ie = int(atoms[idx]), int(iies[idx])
        at = system.atoms[a]
        M = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, system.tmatStore,
                        system.blkSizeTmatStore, ie, system.prel(ie))
        t0 = O.t0(3, int(at.LIZStoreIdx[0]), system.tmatStore, system.blkSizeTmatStore, ie)
        lu, piv, info = sl.lapack.zgetrf(M, overwrite_a=True)
        rhs = np.zeros((M.shape[0], t0.shape[0]), complex)
        rhs[:t0.shape[0]] = t0
        x, info = sl.lapack.zgetrs(lu, piv, rhs)
        with lock:
            refs[idx] = x[:t0.shape[0]].copy()

    pos = [0]

    def take(n):
        out = [order[(pos[0] + q) % len(order)] for q in range(n)]
        pos[0] += n
        return out

    modes = {}
    # (ii) one solve at a time, threaded LAPACK
    with threadpool_limits(limits=ncpu):
        t0 = time.perf_counter()
        n = 0
        while n < 2 or (time.perf_counter() - t0 < budget_s[0] and n < len(order)):
            one(take(1)[0])
            n += 1
        modes["ii_threaded_lapack"] = {"solves_per_s": n / (time.perf_counter() - t0),
                                       "solves": n, "threads": ncpu}
    # (i) one solve per thread, sequential LAPACK (rounds of ncpu solves)
    with threadpool_limits(limits=1), ThreadPoolExecutor(max_workers=ncpu) as ex:
        t0 = time.perf_counter()
        n = 0
        while n < 2 * ncpu or (time.perf_counter() - t0 < budget_s[1] and n < len(order)):
            list(ex.map(one, take(ncpu)))
            n += ncpu
        modes["i_solve_per_thread"] = {"solves_per_s": n / (time.perf_counter() - t0),
                                       "solves": n, "threads": ncpu}
    # one core
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        n = 0
        while n < 2 or (time.perf_counter() - t0 < budget_s[2] and n < len(order)):
            one(take(1)[0])
            n += 1
        modes["one_core"] = {"solves_per_s": n / (time.perf_counter() - t0), "solves": n,
                             "threads": 1}
    return {"ncpu": ncpu, "cpu_model": model, "modes": modes}, refs


def lapack_error(tau, refs):
    """Largest relative (Frobenius) error of tau[idx] against the LAPACK
    tau00 of every sampled solve idx."""
    import numpy as np
    return max(float(np.linalg.norm(tau[i] - r) / np.linalg.norm(r)) for i, r in refs.items())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--natoms", type=int, default=32)
    ap.add_argument("--nenergies", type=int, default=31)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "mixed", "f16", "all"), default="all",
                    help="fp64: plain path only; mixed: also the mixed path with FP32 trailing "
                         "updates; f16: also the mixed path with split-FP16 trailing updates; "
                         "all: both mixed paths.  The fastest mixed path that meets the 1e-10 "
                         "error bar against FP64 is the headline")
    ap.add_argument("--schedule", type=int, default=-1,
                    help="schedule flags (LSMSX_SCHED_*) instead of the autotuner's choice; "
                         "0 = one stream, as used for the rocprofv3 summaries")
    ap.add_argument("--workload", choices=("c2", "c3", "c4", "c5", "ss"), default="c2",
                    help="c2: the headline (fcc 135 sites, mixed vs FP64); c4: uneven 4096-site "
                         "bcc alloy batch (graph + autotuned schedule vs eager); c5: spin-canted "
                         "bcc 137 sites (N=4384), plain tau00 and the adjoint route; ss: "
                         "single-site t-matrices of 64 species x 2 spins at 48 energies")
    ap.add_argument("--profile-step", choices=("fp64", "fp32", "f16"), default=None,
                    help="only run one untimed and one timed eager step of the C2 batch in this "
                         "precision (one stream, no graph, no autotuner, no CPU leg; --schedule "
                         "flags apply, except two streams): the command under rocprofv3 for the "
                         "kernel summaries")
    ap.add_argument("--split-group", type=int, choices=(4, 8, 16), default=8,
                    help="with --profile-step f16: steps per trailing update of the split path")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE readings) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by independent (atom, energy) solves, one batch per GPU)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from lsms_amd import capi
    from lsms_amd import configs as CF

    if not os.path.exists(capi.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()

    if args.workload == "c3":
        return bench_c3(args)
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        return bench_c5(args)
    if args.workload == "ss":
        return bench_ss(args)

    s = CF.config_c2(npts=args.nenergies, natoms=args.natoms)
    ctx = capi.Context(0)
    ctx.load_system(s)
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    prels = np.array([s.prel(i) for i in iies])
    nsolve = len(atoms)
    N = s.nrmat_ns(0)
    k = s.kkrsz_ns

    if args.profile_step:
        prec = {"fp64": capi.PREC_FP64, "fp32": capi.PREC_MIXED_FP32,
                "f16": capi.PREC_MIXED_F16}[args.profile_step]
        ctx.set_graph(False)
        ctx.set_schedule(max(args.schedule, 0) & ~capi.SCHED_TWO_STREAMS)
        ctx.set_split_group(args.split_group)
        ctx.tau00_batched(atoms, iies, prels, k, precision=prec)
        t0 = time.perf_counter()
        tau, info = ctx.tau00_batched(atoms, iies, prels, k, precision=prec)
        el = time.perf_counter() - t0
        iters, _ = ctx.last_stats(nsolve)
        ctx.close()
        print(json.dumps({"profile_step": args.profile_step, "split_group": args.split_group,
                          "solves": nsolve,
                          "ms_per_step": round(1e3 * el, 3), "info_nonzero": int(np.sum(info != 0)),
                          "mean_sweeps": round(float(np.mean(iters)), 3)}), flush=True)
        return

    schedules = {}

    def timed(precision):
        """The library's autotuner picks the schedule for this batch and
        precision (untimed); the timed steps then replay the captured graph."""
        ctx.set_graph(False)
        if args.schedule >= 0:
            ctx.set_schedule(args.schedule)
            schedules[precision] = {"flags": args.schedule, "ms_by_flags": None}
        else:
            best, ms = ctx.autotune(atoms, iies, prels, precision=precision)
            schedules[precision] = {"flags": best, "ms_by_flags": [round(x, 1) for x in ms]}
            if precision == capi.PREC_MIXED_F16:  # the autotuner also picks the step group
                group, gms = ctx.split_group()
                schedules[precision].update(
                    {"split_group": group, "ms_by_split_group": {
                        g: round(float(x), 1) for g, x in zip((4, 8, 16), gms)}})
        ctx.set_graph(True)
        for _ in range(args.warmup):  # the first call captures the graph
            ctx.tau00_batched(atoms, iies, prels, k, precision=precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            tau, info = ctx.tau00_batched(atoms, iies, prels, k, precision=precision)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ctx.set_graph(False)
        assert np.all(info == 0), "singular pivot reported"
        return tau, el

    def kernel_profile(precision):
        """Per-kernel time of one extra step with HIP events on the library's
        stream (eager launches on one stream, so the kernels do not overlap)."""
        ctx.set_profiling(True)
        ctx.reset_profile()
        ctx.tau00_batched(atoms, iies, prels, k, precision=precision)
        ctx.set_profiling(False)
        names = (("assemble", capi.KIND_ASSEMBLE), ("diag", capi.KIND_DIAG),
                 ("panel", capi.KIND_PANEL), ("cross", capi.KIND_CROSS),
                 ("update", capi.KIND_UPDATE),
                 ("last_tile", capi.KIND_FINISH), ("ir_residual", capi.KIND_IR_RESIDUAL),
                 ("ir_solve", capi.KIND_IR_SOLVE))
        return {name: ctx.profile(kind) for name, kind in names}

    tau64, el64 = timed(capi.PREC_FP64)
    fp64_solves = nsolve * args.steps / el64
    prof64 = kernel_profile(capi.PREC_FP64)
    mixed = {}
    runs = {}
    modes = {"mixed": [("fp32_update", capi.PREC_MIXED_FP32)],
             "f16": [("split_f16_update", capi.PREC_MIXED_F16)],
             "all": [("fp32_update", capi.PREC_MIXED_FP32),
                     ("split_f16_update", capi.PREC_MIXED_F16)]}.get(args.precision, [])
    for name, mode in modes:
        tm, elm = timed(mode)
        iters, _ = ctx.last_stats(nsolve)
        profm = kernel_profile(mode)
        errm = max(float(np.linalg.norm(tm[n] - tau64[n]) / np.linalg.norm(tau64[n]))
                   for n in range(nsolve))
        runs[name] = (tm, elm, profm)
        mixed[name] = {"solves_per_s": round(nsolve * args.steps / elm, 3),
                       "max_rel_err_vs_fp64": errm,
                       "refinement_sweeps": {"min": int(iters.min()), "max": int(iters.max()),
                                             "mean": round(float(iters.mean()), 3),
                                             "fp64_fallbacks": int(np.sum(iters < 0))},
                       "kernel_ms_per_step": {kk: round(v[0], 3) for kk, v in profm.items()}}
    ctx.close()

    # CPU baseline, and LAPACK tau00 of every solve it timed: the accuracy gate
    cpu = None
    lapack_err = {}
    cpu_info = None
    if not args.no_cpu:
        cpu_info, refs = cpu_baseline(s, atoms, iies)
        lapack_err["fp64"] = lapack_error(tau64, refs)
        for name in runs:
            lapack_err[name] = lapack_error(runs[name][0], refs)
            mixed[name]["max_rel_err_vs_lapack"] = lapack_err[name]
        best_mode = max(cpu_info["modes"], key=lambda m: cpu_info["modes"][m]["solves_per_s"])
        bm = cpu_info["modes"][best_mode]
        cpu = {"value": round(bm["solves_per_s"], 4), "unit": "solves/s", "cores": bm["threads"],
               "kind": "port",
               "sample": (f"oracle C assembly + LAPACK zgetrf/zgetrs (scipy OpenBLAS) on "
                          f"{len(refs)} of the {nsolve} solves (N={N}); best of three modes "
                          f"({best_mode}); host: {cpu_info['ncpu']} CPUs, {cpu_info['cpu_model']}"),
               "modes": {m: {"solves_per_s": round(v["solves_per_s"], 4), "solves": v["solves"],
                             "threads": v["threads"]} for m, v in cpu_info["modes"].items()},
               "one_core_solves_per_s": round(cpu_info["modes"]["one_core"]["solves_per_s"], 4),
               "host_cpus": cpu_info["ncpu"]}
    # headline: the fastest mixed path within 1e-10 of the FP64 path on every
    # solve and of LAPACK on every sampled solve
    headline = "fp64"
    tau, el, prof = tau64, el64, prof64
    for name, (tm, elm, profm) in runs.items():
        ok = mixed[name]["max_rel_err_vs_fp64"] <= 1e-10 and lapack_err.get(name, 0.0) <= 1e-10
        mixed[name]["meets_1e-10"] = ok
        if ok and elm < el:
            headline = name
            tau, el, prof = tm, elm, profm
    gate_ok = lapack_err.get(headline, 0.0) <= 1e-10
    ms_per_step = 1e3 * el / args.steps
    solves_per_s = nsolve * args.steps / el
    upd_ms, upd_n = prof["update"]
    nt = (N + 63) // 64
    # (kernel, MFMA peak, executed real flops per algorithmic flop, C and panel
    # bytes per complex element, dtype, precision)
    # (kernel, elimination steps per trailing update)
    kernel, peak, flop_mult, cb, pb, dtype, prec, group = {
        # 3M: 3 real products (6 flops) per complex multiply-add (8 flops)
        "fp64": ("k_lu_trail<double>", FP64_PEAK_TFLOPS, 0.75, 16, 16, "f64", "FP64", 2),
        "fp32_update": ("k_lu_trail<float>", FP32_PEAK_TFLOPS, 0.75, 8, 8,
                        "c32-factor+f64-refine",
                        "mixed: complex FP32 LU on MFMA + FP64 iterative refinement", 2),
        # 4M x 3 split products: 12 real products (24 flops) per complex multiply-add
        "split_f16_update": ("k_lu_trail_h", FP16_PEAK_TFLOPS, 3.0, 8, 8,
                             "f16x2-split-update+f32-factor+f64-refine",
                             "mixed: complex FP32 LU with split-FP16 trailing updates on the FP16 "
                             "matrix cores + FP64 iterative refinement",
                             schedules.get(capi.PREC_MIXED_F16, {}).get("split_group", 8)),
    }[headline]
    upd_flops = update_kernel_flops([nt] * nsolve, group)
    upd_bytes = update_kernel_bytes([nt] * nsolve, cb, pb, group)
    t_s = upd_ms * 1e-3 if upd_ms > 0 else float("nan")
    achieved_tf = upd_flops / t_s / 1e12
    achieved_gbs = upd_bytes / t_s / 1e9
    # the bound is the roofline of the algorithmic work: algorithmic flops at
    # the MFMA peak of the kernel's arithmetic against algorithmic bytes at
    # the HBM peak (the split-FP16 kernel executes 3x its algorithmic flops and
    # sits at the ridge point in executed terms; that utilisation is reported
    # as mfma_frac_executed)
    t_mfma = upd_flops / (peak * 1e12)
    t_hbm = upd_bytes / (HBM_PEAK_GBS * 1e9)
    bound = "mfma" if t_mfma >= t_hbm else "hbm"
    # the counter figures in profiles/ are of this workload
    c2 = (args.natoms, args.nenergies) == (32, 31)
    # the split kernel's figures are recorded per step group (k_lu_trail_h@8 ...)
    pkey = f"{kernel}@{group}" if headline == "split_f16_update" else kernel
    prof_flops = profiled_value("mfma_flops_executed_per_launch", pkey) if c2 else None
    t_launch = t_s / max(upd_n, 1)
    path_tf = solve_flops(N, k) * solves_per_s / 1e12

    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": f"fcc a=6.82 bohr, LIZ r=13.0 (135 sites, N={N}), lmax=3, "
                        f"{na} atoms x {ne} contour energies, batched",
            "solves_per_step": nsolve, "N": N, "kkrsz_ns": k,
            "precision": prec,
        },
        "roofline": {
            "bound": bound, "kernel": kernel,
            "achieved": round(achieved_tf, 3) if bound == "mfma" else round(achieved_gbs, 1),
            "peak": peak if bound == "mfma" else HBM_PEAK_GBS,
            "unit": "TFLOP/s" if bound == "mfma" else "GB/s",
            "frac": round(achieved_tf / peak, 4) if bound == "mfma" else round(achieved_gbs / HBM_PEAK_GBS, 4),
            "algorithmic_flops_per_launch": upd_flops / max(upd_n, 1),
            "algorithmic_bytes_per_launch": upd_bytes / max(upd_n, 1),
            # executed MFMA flops per launch read from SQ_INSTS_VALU_MFMA_MOPS
            # (profile of this workload) over the live launch time
            "mfma_frac_executed": (round(prof_flops / t_launch / 1e12 / peak, 4) if prof_flops
                                   else round(achieved_tf * flop_mult / peak, 4)),
            "mfma_frac_executed_from": "counter" if prof_flops else "flop count",
            "mfma_util_counter": profiled_value("mfma_util_counter", pkey) if c2 else None,
            "hbm_frac_algorithmic": round(achieved_gbs / HBM_PEAK_GBS, 4),
            "traffic": profiled_value("traffic_bytes_per_launch", pkey) if c2 else None,
            "avg_launch_ms": round(upd_ms / max(upd_n, 1), 4), "launches": upd_n,
        },
        # algorithmic FP64-equivalent flops of the whole path (8/3 N^3 + 8 N^2 k per
        # solve) at the headline rate, against the FP64 matrix peak; on a mixed
        # headline they run on the FP32/FP16 matrix cores, so the ratio can exceed 1
        "path_roofline": {"achieved": round(path_tf, 3), "unit": "TFLOP/s",
                          "fp64_equivalent_flops_vs_fp64_peak": round(path_tf / FP64_PEAK_TFLOPS, 4),
                          "flops_per_solve": solve_flops(N, k)},
        "fp64_path": {"solves_per_s": round(fp64_solves, 3),
                      "frac_of_fp64_peak": round(solve_flops(N, k) * fp64_solves / 1e12
                                                 / FP64_PEAK_TFLOPS, 4),
                      "max_rel_err_vs_lapack": lapack_err.get("fp64"),
                      "kernel_ms_per_step": {kk: round(v[0], 3) for kk, v in prof64.items()}},
        "mixed_paths": mixed or None,
        "schedules": {{0: "fp64", 1: "fp32_update", 2: "split_f16_update"}[p]: v
                      for p, v in schedules.items()},
        "speedup_mixed_vs_fp64": round(solves_per_s / fp64_solves, 3) if headline != "fp64" else None,
        "max_rel_err_vs_lapack": lapack_err.get(headline),
        "accuracy_gate": ("LAPACK, sampled solves" if lapack_err else "FP64 path only (--no-cpu)"),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    if not gate_ok:
        print(f"headline path {headline} is more than 1e-10 off LAPACK "
              f"({lapack_err[headline]:.3e})", file=sys.stderr)
        sys.exit(1)


def _timed_runs(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t0) / steps


def bench_c3(args):
    """48 energies (32 contour points + 16 at Im E = 0.001-0.01 Ry) for the
    32 fcc-135 atoms: mixed vs FP64 throughput, the per-energy tau00 error of
    the mixed path, the multiple-scattering IDOS error (unit radial
    integrals) and the Lloyd IDOS of the FP64 adjoint route."""
    import numpy as np
    from lsms_amd import capi
    from lsms_amd import configs as CF
    from lsms_amd import contour
    from oracle import oracle as O
    s = CF.config_c3()
    ctx = capi.Context(0)
    ctx.load_system(s)
    ctx.upload_dtmats(s.dtmat_store(), s.blkSizeTmatStore, len(s.energies))
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    prels = np.array([s.prel(i) for i in iies])
    (t64, i64), el64 = _timed_runs(lambda: ctx.tau00_batched(atoms, iies, prels, 16),
                                   args.steps, args.warmup)
    t0s = [np.array([O.t0(3, int(s.atoms[a].LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, ie)
                     for ie in range(ne)]) for a in range(na)]
    n64 = [contour.idos_ms(t64[a * ne:(a + 1) * ne], t0s[a], s.weights) for a in range(na)]
    paths = {}
    for name, mode in (("fp32_update", capi.PREC_MIXED_FP32),
                       ("split_f16_update", capi.PREC_MIXED_F16)):
        run = lambda: ctx.tau00_batched(atoms, iies, prels, 16, precision=mode)
        (tm, im), elm = _timed_runs(run, args.steps, args.warmup)
        iters, _ = ctx.last_stats(len(atoms))
        assert np.all(im == 0)
        err = np.array([np.linalg.norm(tm[n] - t64[n]) / np.linalg.norm(t64[n])
                        for n in range(len(atoms))]).reshape(na, ne)
        idos_err = max(abs(contour.idos_ms(tm[a * ne:(a + 1) * ne], t0s[a], s.weights) - n64[a])
                       / abs(n64[a]) for a in range(na))
        paths[name] = {
            "solves_per_s": round(len(atoms) / elm, 3), "ms_per_step": round(1e3 * elm, 3),
            "speedup_vs_fp64": round(el64 / elm, 3),
            "per_energy_max_rel_err": [float("%.3g" % x) for x in err.max(axis=0)],
            "refinement_sweeps": {"min": int(iters.min()), "max": int(iters.max()),
                                  "mean": round(float(iters.mean()), 3),
                                  "fp64_fallbacks": int(np.sum(iters < 0))},
            "idos_ms_max_rel_err_vs_fp64": float("%.3g" % idos_err)}
    on = np.nonzero(np.abs(s.weights) > 0)[0]
    la, ll, _, ia = ctx.tau00_adjoint(np.repeat(np.arange(na), len(on)), np.tile(on, na),
                                      np.tile(s.energies[on], na), 16)
    ctx.close()
    assert np.all(i64 == 0) and np.all(ia == 0)
    idos_l = [contour.idos_lloyd(ll[a * len(on):(a + 1) * len(on)], s.weights[on]) for a in range(na)]
    best = max(paths, key=lambda k: paths[k]["solves_per_s"])
    out = {
        "metric": "tau00 solves/sec at lmax=3 LIZ=135 with 16 near-axis energies: mixed vs FP64",
        "value": paths[best]["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": paths[best]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "mixed (" + best + ") + f64 refinement", "data": "synthetic",
        "config": {"workload": f"fcc a=6.82, LIZ r=13.0 (135 sites, N=2160), {na} atoms x "
                               f"{ne} energies (16 at Im E 0.001..0.01 Ry)",
                   "solves_per_step": len(atoms)},
        "fp64_solves_per_s": round(len(atoms) / el64, 3),
        "mixed_paths": paths,
        "energies": [[round(float(e.real), 5), round(float(e.imag), 5)] for e in s.energies],
        "idos_ms_fp64": [round(x, 10) for x in n64],
        "idos_lloyd_fp64": [round(x, 8) for x in idos_l],
    }
    print(json.dumps(out))


def bench_c4(args):
    """Uneven batch: 64 alloy atoms x 32 energies, KKR sizes from the LIZ
    radii {11.5, 13.3} bohr with 10% vacancies; eager default schedule
    against hipGraph replay with the autotuned schedule (FP64)."""
    import numpy as np
    from lsms_amd import capi
    from lsms_amd import configs as CF
    s = CF.config_c4()
    ctx = capi.Context(0)
    ctx.load_system(s)
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    prels = np.array([s.prel(i) for i in iies])
    ns = [s.nrmat_ns(a) for a in atoms]
    flops = sum(solve_flops(n, 16) for n in ns)
    run = lambda: ctx.tau00_batched(atoms, iies, prels, 16)
    ctx.set_schedule(0)
    (tau0, info0), t_eager = _timed_runs(run, args.steps, args.warmup)
    best, ms = ctx.autotune(atoms, iies, prels)
    ctx.set_graph(True)
    (tau1, info1), t_tuned = _timed_runs(run, args.steps, args.warmup)
    ctx.set_graph(False)
    ctx.close()
    assert np.all(info0 == 0) and np.all(info1 == 0)
    out = {
        "metric": "tau00 solves/sec, uneven batch (bcc alloy, 4096 sites, 10% vacancies)",
        "value": round(len(atoms) / t_tuned, 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * t_tuned, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"bcc a=5.42, 16x16x8 cells, 10% vacancies, {na} atoms x {ne} "
                               f"energies, LIZ r in {{11.5, 13.3}}, lmax=3",
                   "solves_per_step": len(atoms), "N_min": int(min(ns)), "N_max": int(max(ns))},
        "eager_default_schedule_solves_per_s": round(len(atoms) / t_eager, 3),
        "graph_autotuned_solves_per_s": round(len(atoms) / t_tuned, 3),
        "autotune": {"best_flags": best, "ms_by_flags": [round(x, 2) for x in ms],
                     "flags": "1 right-aligned steps, 2 XCD swizzle"},
        "bitwise_equal_to_eager": bool(np.array_equal(tau0, tau1)),
        "path_roofline": {"achieved": round(flops / t_tuned / 1e12, 3), "unit": "TFLOP/s",
                          "frac_of_fp64_peak": round(flops / t_tuned / 1e12 / FP64_PEAK_TFLOPS, 4)},
    }
    print(json.dumps(out))


def bench_c5(args):
    """Spin-canted bcc, 137-site LIZ (N = 4384, kkrsz_ns = 32), 8 atoms x 31
    energies: the plain FP64 tau00 batch and the adjoint route (explicit
    inverse, Lloyd term and the t-store gradient of a random cotangent)."""
    import numpy as np
    from lsms_amd import capi
    from lsms_amd import configs as CF
    s = CF.config_c5()
    ctx = capi.Context(0)
    ctx.load_system(s)
    ctx.upload_dtmats(s.dtmat_store(), s.blkSizeTmatStore, len(s.energies))
    na, ne = len(s.atoms), len(s.energies)
    atoms = np.repeat(np.arange(na), ne)
    iies = np.tile(np.arange(ne), na)
    prels = np.array([s.prel(i) for i in iies])
    N = s.nrmat_ns(0)
    rng = np.random.default_rng(0)
    V = rng.normal(size=(len(atoms), 32, 32)) + 1j * rng.normal(size=(len(atoms), 32, 32))
    (tau, info), t_plain = _timed_runs(lambda: ctx.tau00_batched(atoms, iies, prels, 32),
                                       args.steps, args.warmup)
    adj = lambda: ctx.tau00_adjoint(atoms, iies, s.energies[iies], 32, lloyd=True, tau00_bar=V,
                                    store_shape=s.tmatStore.shape)
    (tau_a, ll, grad, info_a), t_adj = _timed_runs(adj, max(1, args.steps // 2), 1)
    ctx.close()
    assert np.all(info == 0) and np.all(info_a == 0)
    err = max(float(np.linalg.norm(tau_a[b] - tau[b]) / np.linalg.norm(tau[b])) for b in range(len(atoms)))
    fl = solve_flops(N, 32)
    Ns = (N + 63) // 64 * 64
    inv_fl = 8.0 * Ns ** 3 * (1.0 / 3 + 1.0 / 6 + 1.0 / 2)
    out = {
        "metric": "tau00 solves/sec, spin-canted bcc LIZ=137 (N=4384); adjoint route",
        "value": round(len(atoms) / t_plain, 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * t_plain, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"bcc a=5.42, n_spin_cant=2, LIZ r=13.3 (137 sites, N={N}), "
                               f"{na} atoms x {ne} energies", "solves_per_step": len(atoms)},
        "path_roofline": {"achieved": round(fl * len(atoms) / t_plain / 1e12, 3), "unit": "TFLOP/s",
                          "frac_of_fp64_peak": round(fl * len(atoms) / t_plain / 1e12 / FP64_PEAK_TFLOPS, 4)},
        "adjoint": {"solves_per_s": round(len(atoms) / t_adj, 3),
                    "ms_per_step": round(1e3 * t_adj, 3),
                    "work": "assembly + LU + explicit inverse + Lloyd trace + t-store gradient",
                    "inverse_tflops": round(inv_fl * len(atoms) / t_adj / 1e12, 3),
                    "max_rel_diff_tau00_vs_plain": err},
    }
    print(json.dumps(out))


def bench_ss(args):
    """Single-site scattering (single_scatterer_nonrel.f via semrel_c.f) of
    64 species x 2 spins = 128 potentials (jmt = 450, iprpts = 1000) at 48
    contour energies, lmax = 3, scalar-relativistic: one call per step.
    Reports (potential, energy) solves/s for t-matrices only (what the tau00
    path consumes) and with the radial functions zlr, jlr copied back."""
    import numpy as np
    from lsms_amd import capi
    from lsms_amd import configs as CF
    rr, vr, jm, jw, rs, e, pr = CF.single_site_batch(nspecies=64, ne=48)
    ctx = capi.Context(0)
    run_t = lambda: ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr, radial=False)
    run_r = lambda: ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr, radial=True)
    (m, _, _, _, rc), t_t = _timed_runs(run_t, args.steps, args.warmup)
    (_, tm, zl, jl, rc2), t_r = _timed_runs(run_r, max(1, args.steps // 2), 1)
    assert rc == 0 and rc2 == 0
    npair = len(rr) * len(e)
    # muffin-tin DOS and green(r) of every (potential, energy) with tau00 = t
    # (the single-site limit), radial functions from the call above
    pot_of = np.tile(np.arange(len(rr)), len(e))
    gf = lambda: ctx.green_mt(3, tm.reshape(npair, 16, 16), zl.reshape(npair, 4, -1),
                              jl.reshape(npair, 4, -1), pot_of, rr, jm, jw, rs)
    _, t_g = _timed_runs(gf, max(1, args.steps // 2), 1)
    # the same chain with the radial functions kept in HBM: single-site solve
    # (radial functions computed, not copied) then the Green function
    ctx.set_single_site_resident(True)
    tau_t = tm.reshape(npair, 16, 16)
    pairs = np.arange(npair)

    def chain():
        ctx.single_site_nonrel(0, 3, vr, rr, jm, jw, rs, e, pr, radial=False)
        return ctx.green_mt_resident(tau_t, pairs, rr.shape[1], with_green=False)
    _, t_c = _timed_runs(chain, max(1, args.steps // 2), 1)
    ctx.close()
    out = {
        "metric": "single-site solves/sec (potential x energy, lmax=3, scalar-relativistic)",
        "value": round(npair / t_t, 1), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * t_t, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(rr)} model potentials (64 species x 2 spins, Z 22-29, "
                               f"jmt=450, iprpts=1000) x {len(e)} energies",
                   "solves_per_step": npair},
        "with_radial_functions": {"solves_per_s": round(npair / t_r, 1),
                                  "ms_per_step": round(1e3 * t_r, 3)},
        "green_mt": {"solves_per_s": round(npair / t_g, 1), "ms_per_call": round(1e3 * t_g, 3),
                     "note": "host-array C-ABI call: 0.8 GB of zlr/jlr uploaded per call"},
        "single_site_then_dos_resident": {"solves_per_s": round(npair / t_c, 1),
                                          "ms_per_step": round(1e3 * t_c, 3)},
    }
    if not args.no_cpu:
        import time as _t
        from oracle import oracle as O
        t0 = _t.perf_counter()
        n = 0
        while _t.perf_counter() - t0 < 3.0:
            q, k = n % len(rr), (7 * n) % len(e)
            O.single_scatterer_nonrel(0, 3, e[k], pr[k], vr[q], rr[q], jm[q], jw[q], rs[q])
            n += 1
        dt = _t.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n / dt, 1), "unit": "solves/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{n} (potential, energy) solves of the same batch "
                                         "through the oracle's restatement, one thread"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
