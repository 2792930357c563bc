#!/usr/bin/env python3
"""Benchmark of the PDLP hot path on one MI355X.

Workload: the 1,000,000-column sparse LP (m 500,000, nnz ~12M, power-law column counts on
[2, 5,000] with mean 12, values N(0,1) x 10^U[-2,2], feasible by construction; the library's
generator, seed 1002), solved to max(rel primal, rel dual, |rel gap|) < 1e-6 with
mixed-precision restarted Halpern PDHG (FP32 values / iterates / accumulation in the steps, FP64
evaluation, restarts and termination; the reference's preconditioner, on the device). One step =
one such solve. --method adaptive gives PDLP with adaptive steps (the enhanced preconditioner).

  value        time to 1e-6 of the mixed-precision solve, in seconds, from the moment the
               problem is resident in HBM (host scaling and upload reported separately as
               setup_s)
  fp64_*       the same LP solved by plain FP64 PDLP (the reference's algorithm) on the GPU
  spmv_*       HBM GB/s of the plain products y = A x and y = A' x (algorithmic bytes
               nnz (s + 4) + 4 (rows + 1) + s cols + s rows)
  roofline     the dominant step kernel of the solve, timed live with HIP events on its stream
  cpu_baseline the reference's own CPU PDLP (oracle/_ref, compiled from its sources) timed on a
               bounded sample of the same LP on one host core, extrapolated to the GPU's
               iteration count

Batched workloads (--workload):
  c4  256 multi-period blending LPs (2.9k-59.5k rows each, 91M nonzeros in all) solved as one
      stacked batch, mixed precision with Halpern steps and the enhanced preconditioner, to 1e-6
      each; value = seconds until the last LP stopped;
      compared with plain FP64 PDLP solving them one at a time (a timed sample of 16, scaled)
  c5  1,024 node relaxations of one knapsack LP (m 20,000, n 50,000, nnz 400k; 1-5% of the
      columns fixed at 0 or 1 per node) solved as one shared-matrix batch (SpMM Halpern steps),
      warm-started from the root LP's answer and primal weight; infeasible nodes detected on the
      device (with the sign-projected Farkas test);
      compared with plain FP64 PDLP solving the nodes one at a time (a stratified sample of 32
      feasible and 16 infeasible nodes, scaled)
  mip a knapsack MIP (m 10, n 60, binary) by branch and bound over waves of 256 batched node
      relaxations (smy_mip_solve); value = seconds to close the gap at 1e-4; compared with one node
      per batch (a 10 s sample of its node rate) and with HiGHS's optimum

Usage: python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--workload c3|c4|c5|mip] [--scale 1.0]
       [--no-fp64] [--no-cpu-baseline] [--no-spmv]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBPS = 8000.0
METRIC = ("time-to-1e-6 rel KKT + SpMV HBM GB/s; FP32/batched vs FP64 PDLP, 1 GPU")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(nnz: int, rows: int, cols: int, s: int) -> int:
    return nnz * (s + 4) + 4 * (rows + 1) + s * cols + s * rows


def step_kernel_bytes(nnz: int, m: int, n: int, s: int, halpern: bool = False) -> dict:
    """Algorithmic HBM bytes of the two fused step kernels (each value / index / offset once,
    the gathered vector once, the per-row and per-column update vectors once). PDHG steps with
    averages: rows of A read 8 and write 5 vectors, rows of A' read 5 and write 3; Halpern steps:
    rows of A read 7 (y, its anchor, the logical variable, its anchor, cost, bounds) and write 2,
    rows of A' read 5 (v, its anchor, cost, bounds) and write 2 (v', 2 v+ - v)."""
    if halpern:
        dual = nnz * (s + 4) + 4 * (m + 1) + s * n + 9 * s * m
        primal = nnz * (s + 4) + 4 * (n + 1) + s * m + 7 * s * n
    else:
        dual = nnz * (s + 4) + 4 * (m + 1) + s * n + 13 * s * m
        primal = nnz * (s + 4) + 4 * (n + 1) + s * m + 8 * s * n
    return {"dual": dual, "primal": primal}


def shared_step_bytes(nnz: int, m: int, n: int, k: int, s: int, coded: bool = False,
                      halpern: bool = False) -> dict:
    """Algorithmic HBM bytes of the shared-matrix step kernels for k instances: the matrix once,
    the gathered iterate block once (k values per row), and the per-row update streams of all
    instances (rows of A: 5 read, 5 written; rows of A': v, average and both bounds read, 3 written;
    with coded bounds one byte per (column, instance) and 4 pairs per column replace the bounds).
    Halpern steps: rows of A read y, its anchor, the logical v and its anchor and write y', v'
    (4 + 2); rows of A' read v, its anchor and the bounds and write w, v' (2 + 2); T(z), written
    at a run's last step only, is left out."""
    bounds = k * n + 8 * s * n if coded else 2 * s * k * n
    if halpern:
        dual = nnz * (s + 4) + 4 * (m + 1) + s * k * n + 6 * s * k * m
        primal = nnz * (s + 4) + 4 * (n + 1) + s * k * m + 4 * s * k * n + bounds
        return {"dual": dual, "primal": primal}
    dual = nnz * (s + 4) + 4 * (m + 1) + s * k * n + 10 * s * k * m
    primal = nnz * (s + 4) + 4 * (n + 1) + s * k * m + 5 * s * k * n + bounds
    return {"dual": dual, "primal": primal}


def shared_bounds_coded(lows, ups) -> bool:
    """Whether the shared-matrix batch keeps these column bounds as codes: at most 4 distinct
    (lower, upper) pairs per column over the instances, padding instances (up to a multiple of 4)
    counting as (0, 0). Pairs are compared by a hash of their bits (a byte model, not a solve)."""
    import numpy as np
    k = len(lows)
    if k == 0 or len(lows[0]) == 0:
        return False
    kp = (k + 3) // 4 * 4
    lo = np.stack([np.asarray(x, dtype=np.float64) for x in lows], axis=1).view(np.int64)
    up = np.stack([np.asarray(x, dtype=np.float64) for x in ups], axis=1).view(np.int64)
    key = lo * np.int64(1000003) ^ up
    if kp > k:
        key = np.concatenate([key, np.zeros((key.shape[0], kp - k), np.int64)], axis=1)
    key = np.sort(key, axis=1)
    distinct = 1 + np.count_nonzero(key[:, 1:] != key[:, :-1], axis=1)
    return bool(distinct.max() <= 4)


def profiled_traffic(workload: str, kernel: str) -> int | None:
    """Calibrated HBM bytes per launch of a step kernel on a workload (c3, c4, c5), from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries: lines `traffic_bytes <workload>
    <kernel> <bytes>` in profiles/*.md. None if no summary has it."""
    for path in sorted((ROOT / "profiles").glob("*.md")):
        for line in path.read_text().splitlines():
            parts = line.split()
            if len(parts) == 4 and parts[0] == "traffic_bytes" and parts[1] == workload and parts[2] == kernel:
                return int(parts[3])
    return None


def spmv_sweep(md) -> dict:
    import numpy as np
    import scipy.sparse as sp
    import torch

    from samaya_amd.device import SpmvPlan

    A = sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))
    out = {}
    for name, M in (("Ax", A.tocsr()), ("Aty", A.T.tocsr())):
        M.sort_indices()
        ptr = torch.tensor(M.indptr.astype(np.int32), device="cuda")
        idx = torch.tensor(M.indices.astype(np.int32), device="cuda")
        for prec, dt, s in (("fp64", torch.float64, 8), ("fp32", torch.float32, 4)):
            val = torch.tensor(M.data, device="cuda", dtype=dt)
            x = torch.rand(M.shape[1], device="cuda", dtype=dt)
            y = torch.zeros(M.shape[0], device="cuda", dtype=dt)
            plan = SpmvPlan(M.indptr.astype(np.int32), ptr, idx, val, prec)
            stream = torch.cuda.current_stream()
            for _ in range(3):
                plan.run(x, y, stream.cuda_stream)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            reps = 20
            e0.record(stream)
            for _ in range(reps):
                plan.run(x, y, stream.cuda_stream)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            b = spmv_bytes(M.nnz, M.shape[0], M.shape[1], s)
            out[f"{name}_{prec}"] = {"us": round(ms * 1e3, 2), "gbps": round(b / ms / 1e6, 1)}
            del val, x, y, plan
    return out


def count_order(md, algorithm: str, column_order: str = "auto"):
    """The column order of the solver's model-level solve (PdlpOptions.column_order): "degree" by
    decreasing count; "natural" None (the model's order); "touch" the hot columns (count >= 2x the
    mean) by count, then the others by the first row holding them; "auto" the touch order for the
    enhanced algorithm when the longest column is at least 16x the mean, else None."""
    counts = np.diff(md.start)
    if column_order == "natural" or md.n == 0:
        return None
    if column_order == "auto" and (algorithm != "enhanced" or counts.max() < 16 * max(md.nnz / md.n, 1.0)):
        return None
    if column_order in ("touch", "auto"):
        idx = np.asarray(md.index)
        first = np.full(md.n, md.m, dtype=np.int64)
        nonempty = counts > 0
        first[nonempty] = np.minimum.reduceat(idx, np.asarray(md.start)[:-1][nonempty])
        hot = counts >= 2.0 * (md.nnz / md.n)
        hot_idx, cold_idx = np.flatnonzero(hot), np.flatnonzero(~hot)
        return np.concatenate([hot_idx[np.argsort(-counts[hot_idx], kind="stable")],
                               cold_idx[np.argsort(first[cold_idx], kind="stable")]])
    return np.argsort(-counts, kind="stable")


def kernel_roofline(md, precision: str, adaptive: bool = False, acc64: bool = False,
                    order=None, halpern: float | None = None) -> dict:
    """Live per-kernel timing of the two step kernels on the workload's matrix, with its columns
    in `order` when the solve uses one (the adaptive attempt's kernels when `adaptive`: same byte
    model, plus the per-attempt decision kernel; `acc64`: the FP32 kernels that accumulate in
    FP64)."""
    import numpy as np
    import scipy.sparse as sp

    from samaya_amd.device import PdhgDevice

    A = sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))
    obj, col_lower, col_upper = md.obj, md.col_lower, md.col_upper
    if order is not None:
        A = A[:, order]
        A.sort_indices()
        obj, col_lower, col_upper = obj[order], col_lower[order], col_upper[order]
    T = A.T.tocsc()
    T.sort_indices()
    nt = md.n + md.m
    lower = np.concatenate([col_lower, md.row_lower])
    upper = np.concatenate([col_upper, md.row_upper])
    cost = np.concatenate([obj, np.zeros(md.m)])
    dev = PdhgDevice.create((md.m, md.n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data),
                            (md.n, md.m, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data),
                            cost, lower, upper, "fp32" if precision == "fp32" else "fp64")
    if acc64:
        dev.set_fp64_accumulation(True)
    if halpern is not None:
        dev.set_halpern(True, halpern)
        adaptive = False
    decide_ms = None
    if adaptive:
        dual_ms, primal_ms, decide_ms = dev.time_adaptive_kernels(50)
    else:
        dual_ms, primal_ms = dev.time_kernels(50)
    s = 4 if precision == "fp32" else 8
    b = step_kernel_bytes(md.nnz, md.m, md.n, s, halpern is not None)
    ops = (("HalpernRowsOp", "HalpernColsOp") if halpern is not None else
           ("AdaptDualOp", "AdaptPrimalOp") if adaptive else ("DualRowsOp", "PrimalColsOp"))
    return {"dual_ms": dual_ms, "primal_ms": primal_ms, "decide_ms": decide_ms, "bytes": b,
            "ops": {"dual": ops[0], "primal": ops[1]}}


def host_info() -> dict:
    """The host cores the CPU baseline ran on: usable CPUs, the model name, the flags the oracle
    is built with (oracle/Makefile, as the reference's own: -O3 -ffp-contract=off)."""
    model = ""
    try:
        model = next((ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo")
                      if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"nproc": len(os.sched_getaffinity(0)), "cpu_model": model,
            "flags": "-std=c++20 -O3 -ffp-contract=off"}


def calibration() -> dict | None:
    """tests/golden/cpu_calibration.json: the reference's CPU PDLP over the oracle restatement,
    time ratio on one core (the 2,000 x 5,000 LP, measured where the reference tree is)."""
    path = ROOT / "tests" / "golden" / "cpu_calibration.json"
    return json.loads(path.read_text()) if path.exists() else None


def port_rate(md, its) -> tuple[list, float, float]:
    """The oracle restatement of the reference's CPU PDLP (one core, FP64, parity mode) for two
    short runs of `its` iterations: (times, seconds per iteration, setup seconds)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    m = O.M.from_product(md)
    times = []
    for k in its:
        t0 = time.perf_counter()
        O.oracle_solve_lp(m, 1e-6, k)
        times.append(time.perf_counter() - t0)
    per_it = (times[1] - times[0]) / (its[1] - its[0])
    return times, per_it, times[0] - per_it * its[0]


def cpu_baseline(md, gpu_iterations: int, sample_iters=(64, 576)) -> dict:
    """The CPU baseline of one LP: the oracle restatement timed for two short runs on one host
    core; the difference gives the time per iteration, extrapolated to the GPU FP64 run's
    iteration count. `reference_equivalent` scales it by the committed reference-over-restatement
    ratio."""
    times, per_it, setup = port_rate(md, sample_iters)
    value = setup + per_it * gpu_iterations
    cal = calibration()
    out = {"value": round(value, 2), "unit": "s", "cores": 1, "kind": "port",
           "sample": (f"{sample_iters[0]} and {sample_iters[1]} FP64 PDLP iterations of the same "
                      f"LP on one core ({times[0]:.1f} s, {times[1]:.1f} s): "
                      f"{per_it * 1e3:.1f} ms/iteration, setup {setup:.1f} s, extrapolated to "
                      f"{gpu_iterations} iterations"), **host_info()}
    if cal is not None:
        out["reference_over_port"] = round(cal["reference_over_port"], 4)
        out["reference_equivalent"] = round(value * cal["reference_over_port"], 2)
    return out


def copy_calibration() -> int:
    """One plain device-to-device copy of 1 GiB (HIP's own copy kernel). Under rocprofv3 --pmc its
    FETCH_SIZE and WRITE_SIZE give the counters' scale: measured 524315 KiB and 1048576 KiB, so
    fetches are multiplied by 2 and writes by 1 to get bytes."""
    import torch
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {a.numel() * 8} bytes")
    return 0


BUDGET_S = 480.0  # the default run ends well inside 10 minutes: sub-results past this are skipped
SUB_COST_S = {"c4": 110.0, "c5": 60.0, "mip": 30.0}  # measured cost of each sub-result, with margin


def relerr(a, b) -> float | None:
    return None if a is None or b is None else abs(a - b) / (1 + abs(b))


def main() -> int:
    t_begin = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c3", choices=["c3", "c4", "c5", "mip"])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-spmv", action="store_true")
    ap.add_argument("--no-batches", action="store_true",
                    help="the default (c3) line without its c4 / c5 / mip sub-results")
    ap.add_argument("--polish-rounds", type=int, default=0,
                    help="c3: also time one solve with this many FP64 polish rounds (the "
                         "continuation at tol / 10 after a verifier rejection)")
    ap.add_argument("--method", default="halpern", choices=["halpern", "adaptive", "constant"],
                    help="the single LP's iteration: restarted Halpern PDHG with reflection "
                         "(default), or PDLP with adaptive / constant steps")
    ap.add_argument("--algorithm", default=None, choices=["reference", "enhanced"],
                    help="preconditioner: the reference's 2-norm variant or Ruiz + Pock-Chambolle "
                         "(default: reference for Halpern steps, enhanced otherwise)")
    ap.add_argument("--fp64-accumulation", action="store_true",
                    help="mixed precision: FP32 steps sum their products in FP64 (the default line "
                         "is specified with FP32 accumulation)")
    ap.add_argument("--adaptive-steps", default="auto", choices=["auto", "on", "off"],
                    help="PDLP's adaptive step sizes for the mixed-precision solves (auto: on for "
                         "the single LP and the stacked batch; the shared-matrix batch has none)")
    ap.add_argument("--c4-method", default="halpern", choices=["halpern", "adaptive", "constant"],
                    help="the blending batch's iteration (with the enhanced preconditioner)")
    ap.add_argument("--c5-method", default="halpern", choices=["halpern", "constant"],
                    help="the node-relaxation batch's iteration (shared-matrix batch)")
    ap.add_argument("--reflection", type=float, default=1.0, help="Halpern steps: the reflection")
    ap.add_argument("--column-order", default=None, choices=["auto", "natural", "degree", "touch"],
                    help="the single LP's column order (default: touch for Halpern steps: hot "
                         "columns by count, the rest by first row; auto otherwise, which takes the "
                         "touch order for the enhanced algorithm when the counts are skewed)")
    ap.add_argument("--pid-weight", action="store_true",
                    help="the primal weight by a PI controller on log(dy / dx) at restarts")
    ap.add_argument("--verbose", action="store_true", help="progress lines of long batches (stdout)")
    ap.add_argument("--fp32-switch-tol", type=float, default=None,
                    help="mixed precision: FP32 steps until this distance to termination (or a "
                         "stall); default the tolerance for the single LP's Halpern steps and the "
                         "node relaxations, 1e-5 otherwise")
    ap.add_argument("--fp32-stall-level", type=float, default=None,
                    help="mixed precision: a plateau of the FP32 steps below this distance switches "
                         "to FP64 (0: never; default 0 for the single LP's Halpern steps, 1e-4 "
                         "otherwise)")
    ap.add_argument("--max-iterations", type=int, default=-1,
                    help="cap each solve (profiling runs); the line then reports no speedup")
    ap.add_argument("--kernels-only", action="store_true",
                    help="--workload c4 / c5: time the batch's step kernels and exit, without the "
                         "solve (for counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device and exit: the reference transfer that "
                         "scales rocprofv3's FETCH_SIZE / WRITE_SIZE counts to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the batched and "
              "single-LP shardings are designed but not built); run with --gpus 1",
              file=sys.stderr)
        return 2
    on = args.adaptive_steps == "on"
    auto = args.adaptive_steps == "auto"
    # The single LP (c3 line): its iteration, preconditioner, column order and precision rule.
    args.halpern = args.method == "halpern"
    args.adaptive = args.method == "adaptive"
    args.c3_algorithm = args.algorithm or ("reference" if args.halpern else "enhanced")
    args.c3_column_order = args.column_order or ("touch" if args.halpern else "auto")
    args.c3_switch_tol = args.fp32_switch_tol if args.fp32_switch_tol is not None else (
        args.tol if args.halpern else 1e-5)
    args.c3_stall_level = args.fp32_stall_level if args.fp32_stall_level is not None else (
        0.0 if args.halpern else 1e-4)
    # The batches (c4 / c5 lines and sub-results).
    args.halpern_c4 = args.c4_method == "halpern"
    args.adaptive_c4 = args.c4_method == "adaptive" or (args.c4_method == "constant" and on)
    args.fp32_switch_tol_c5 = args.fp32_switch_tol if args.fp32_switch_tol is not None else 1e-6
    if args.fp32_switch_tol is None:
        args.fp32_switch_tol = 1e-6 if args.workload == "c5" else 1e-5
    if args.algorithm is None:
        args.algorithm = "enhanced"
    args.sub_steps = 1

    import torch

    import samaya_amd as smy

    assert torch.cuda.is_available() and smy.gpu_available(), "no usable MI355X"
    torch.cuda.set_device(0)
    if args.workload in ("c4", "c5"):
        return batch_line(args, args.workload)
    if args.workload == "mip":
        return mip_line(args)
    t0 = time.perf_counter()
    md = smy.generators.c2(args.scale)
    log(f"LP m={md.m} n={md.n} nnz={md.nnz} generated in {time.perf_counter() - t0:.1f} s")
    nm = md.native()
    opts = smy.PdlpOptions(tol=args.tol, precision="mixed", algorithm=args.c3_algorithm, detect=False,
                           verify=True, polish_rounds=0, max_iterations=args.max_iterations,
                           fp32_switch_tol=args.c3_switch_tol, adaptive_steps=args.adaptive,
                           fp64_accumulation=args.fp64_accumulation, halpern=args.halpern,
                           reflection=args.reflection, pid_primal_weight=args.pid_weight,
                           column_order=args.c3_column_order, fp32_stall_level=args.c3_stall_level)

    for w in range(args.warmup):
        r = smy.solve_lp(nm, opts)
        log(f"warmup {w}: {r.status_name} {r.iterations} it in {r.solve_seconds:.2f} s")

    torch.cuda.synchronize()
    t_start = time.perf_counter()
    runs, walls = [], []
    for k in range(args.steps):
        t1 = time.perf_counter()
        r = smy.solve_lp(nm, opts)
        walls.append(time.perf_counter() - t1)
        runs.append(r)
        log(f"step {k}: {r.status_name} {r.iterations} it ({r.fp32_iterations} FP32), "
            f"setup {r.setup_seconds:.2f} s, iterate {r.solve_seconds - r.setup_seconds:.3f} s, "
            f"verified {r.verified} ({r.max_primal_violation:.1e}, {r.max_dual_violation:.1e}), "
            f"objective {r.objective!r}, dual bound {r.dual_bound!r}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    iterate_s = sum(r.solve_seconds - r.setup_seconds for r in runs) / len(runs)
    setup_s = sum(r.setup_seconds for r in runs) / len(runs)
    last = runs[-1]
    capped = args.max_iterations > 0
    assert capped or all(r.status == smy.OPTIMAL for r in runs), [r.message for r in runs]

    fp64 = None
    if not args.no_fp64 and not capped:
        # Plain FP64 PDLP: the reference's algorithm and precision on the same GPU, verified too.
        fp64 = smy.solve_lp(nm, smy.PdlpOptions(tol=args.tol, precision="fp64", algorithm="reference",
                                                detect=False, verify=True))
        log(f"fp64: {fp64.status_name} {fp64.iterations} it, iterate "
            f"{fp64.solve_seconds - fp64.setup_seconds:.3f} s, verified {fp64.verified} "
            f"({fp64.max_primal_violation:.1e}, {fp64.max_dual_violation:.1e})")

    polish = None
    if args.polish_rounds > 0 and not capped:
        po = smy.PdlpOptions(**{**opts.__dict__, "polish_rounds": args.polish_rounds})
        t1 = time.perf_counter()
        p = smy.solve_lp(nm, po)
        polish = {"rounds": p.rounds, "iterations": p.iterations, "wall_s": round(time.perf_counter() - t1, 3),
                  "verified": p.verified, "max_primal_violation": p.max_primal_violation,
                  "max_dual_violation": p.max_dual_violation, "objective": p.objective,
                  "time_to_verified_s": round(time.perf_counter() - t1, 3) if p.verified else None}
        log(f"polish {args.polish_rounds}: {p.iterations} it, verified {p.verified} "
            f"({p.max_primal_violation:.1e}, {p.max_dual_violation:.1e})")

    order = count_order(md, args.c3_algorithm, args.c3_column_order)
    rl = kernel_roofline(md, "fp32", args.adaptive, args.fp64_accumulation, order,
                         args.reflection if args.halpern else None)
    dom = "primal" if rl["primal_ms"] >= rl["dual_ms"] else "dual"
    dom_ms = rl[f"{dom}_ms"]
    dom_op = rl["ops"][dom]
    achieved = rl["bytes"][dom] / (dom_ms * 1e-3) / 1e9
    log(f"step kernels fp32: dual {rl['dual_ms'] * 1e3:.1f} us, primal {rl['primal_ms'] * 1e3:.1f} us"
        + (f", decision {rl['decide_ms'] * 1e3:.1f} us" if rl["decide_ms"] is not None else ""))

    spmv = None if args.no_spmv else spmv_sweep(md)
    cpu = None
    if not args.no_cpu_baseline:
        gpu_its = fp64.iterations if fp64 is not None else last.iterations
        cpu = cpu_baseline(md, gpu_its)

    ref = golden("c2_reference") if args.scale == 1.0 else None
    line = {
        "metric": METRIC,
        "value": round(iterate_s, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 1),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp32 steps + fp64 evaluation/polish",
        "method": args.method,
        "algorithm": args.c3_algorithm,
        "adaptive_steps": args.adaptive,
        "halpern": {"reflection": args.reflection} if args.halpern else False,
        "fp32_switch_tol": args.c3_switch_tol,
        "fp32_stall_level": args.c3_stall_level,
        "pid_primal_weight": args.pid_weight,
        "fp64_accumulation": args.fp64_accumulation,
        "detect": False,
        "column_order": ("model" if order is None else
                         "hot by count, then by first row" if args.c3_column_order == "touch" else "by count"),
        "status": last.status_name,
        "data": "synthetic (seeded generator)",
        "config": {"workload": "sparse LP m=500000 n=1000000 nnz~12M power-law columns, "
                               "mixed-precision PDLP to rel KKT 1e-6" if args.scale == 1.0 else
                               f"sparse LP scale {args.scale} (m={md.m}, n={md.n})",
                   "m": md.m, "n": md.n, "nnz": md.nnz, "tol": args.tol},
        "iterations": last.iterations,
        "fp32_iterations": last.fp32_iterations,
        "setup_s": round(setup_s, 3),
        "value_with_setup_s": round(iterate_s + setup_s, 4),
        "solve_lp_wall_s": round(sum(walls) / len(walls), 4),
        "rel_primal_residual": last.primal_residual,
        "rel_dual_residual": last.dual_residual,
        "rel_gap": last.gap,
        "verified": last.verified,
        "max_primal_violation": last.max_primal_violation,
        "max_dual_violation": last.max_dual_violation,
        "time_to_verified_s": None,
        "objective": last.objective,
        "dual_bound": last.dual_bound,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": profiled_traffic("c3", f"{dom_op}<float>") if args.scale == 1.0 else None,
                     "kernel": f"csr_rows_kernel<float, {dom_op}<float>>",
                     "avg_us": round(dom_ms * 1e3, 2), "bytes_per_launch": rl["bytes"][dom]},
        "cpu_baseline": cpu,
    }
    if not last.verified:
        line["verifier_note"] = ("the verifier (1e-6 relative per row and per reduced cost) rejects "
                                 "PDLP answers at rel KKT 1e-6 on this LP, as it rejects the "
                                 "reference's own; FP64 polish rounds do not reach it either")
    if ref is not None:
        line["reference_fp64_objective"] = ref["objective"]
        line["reference_fp64_iterations"] = ref["iterations"]
        line["rel_objective_error_vs_reference_fp64"] = relerr(last.objective, ref["objective"])
        if ref.get("highs_objective") is not None:
            line["highs_objective"] = ref["highs_objective"]
            line["rel_objective_error_vs_highs"] = relerr(last.objective, ref["highs_objective"])
    if fp64 is not None:
        line["fp64_time_to_tol_s"] = round(fp64.solve_seconds - fp64.setup_seconds, 4)
        line["fp64_iterations"] = fp64.iterations
        line["fp64_objective"] = fp64.objective
        line["fp64_verified"] = fp64.verified
        line["fp64_max_primal_violation"] = fp64.max_primal_violation
        line["fp64_max_dual_violation"] = fp64.max_dual_violation
        line["speedup_vs_fp64"] = round(line["fp64_time_to_tol_s"] / iterate_s, 3)
        line["rel_objective_error_vs_fp64"] = relerr(last.objective, fp64.objective)
        if ref is not None:
            line["fp64_rel_objective_error_vs_reference_fp64"] = relerr(fp64.objective, ref["objective"])
    if polish is not None:
        line["polish"] = polish
        line["time_to_verified_s"] = polish["time_to_verified_s"]
    if spmv is not None:
        line["spmv"] = spmv
    if not args.no_batches and args.scale == 1.0 and not capped:
        for name, fn in (("c4", c4_result), ("c5", c5_result), ("mip", mip_result)):
            spent = time.perf_counter() - t_begin
            if spent + SUB_COST_S[name] > BUDGET_S:
                line[name] = {"skipped": "budget", "elapsed_s": round(spent, 1)}
                log(f"{name}: skipped ({spent:.0f} s spent of a {BUDGET_S:.0f} s budget)")
                continue
            t1 = time.perf_counter()
            kw = ({"fp64_every": 32, "cpu_probe_every": 4} if name == "c4" else
                  {"feasible_sample": 16, "infeasible_sample": 8} if name == "c5" else {})
            line[name] = fn(args, **kw)
            line[name]["elapsed_s"] = round(time.perf_counter() - t1, 1)
    print(json.dumps(line), flush=True)
    return 0


def base_line(args, value: float, ms_per_step: float, workload: str, extra: dict) -> dict:
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 1),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp32 steps + fp64 evaluation/polish",
        "algorithm": args.algorithm,
        "adaptive_steps": args.adaptive,
        "fp64_accumulation": args.fp64_accumulation,
        "data": "synthetic (seeded generator)",
        "config": {"workload": workload},
    }
    line.update(extra)
    return line


def batch_roofline(dev, nnz: int, m: int, n: int, k: int | None, adaptive: bool = False,
                   coded: bool = False, traffic_of: str | None = None,
                   halpern: float | None = None) -> dict:
    """Live timing of the FP32 step kernels of a batch handle; k: instances of a shared-matrix
    batch (None: stacked batch); adaptive: the adaptive attempt's kernels (stacked batch; same
    byte model); coded: the shared batch keeps its column bounds as codes; traffic_of: the
    workload whose profiled traffic applies (None: the handle is not that full workload)."""
    if halpern is not None:
        dev.set_halpern(True, halpern)
        dual_ms, primal_ms = dev.time_kernels(20)
        ops = ("HalpernRowsOp", "HalpernColsOp")
    elif adaptive:
        dual_ms, primal_ms, _ = dev.time_adaptive_kernels(20)
        ops = ("AdaptDualOp", "AdaptPrimalOp")
    else:
        dual_ms, primal_ms = dev.time_kernels(20)
        ops = ("DualRowsOp", "PrimalColsOp")
    b = (shared_step_bytes(nnz, m, n, k, 4, coded, halpern is not None) if k else
         step_kernel_bytes(nnz, m, n, 4, halpern is not None))
    dom = "primal" if primal_ms >= dual_ms else "dual"
    ms = primal_ms if dom == "primal" else dual_ms
    achieved = b[dom] / (ms * 1e-3) / 1e9
    shared_name = f"shared_halpern_{dom}_kernel" if halpern is not None else f"shared_{dom}_kernel"
    key = f"{shared_name}<float>" if k else f"{ops[1] if dom == 'primal' else ops[0]}<float>"
    kernel = key if k else f"csr_rows_kernel<float, {key}>"
    traffic = profiled_traffic(traffic_of, key) if traffic_of else None
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": traffic, "kernel": kernel,
            "avg_us": round(ms * 1e3, 2), "bytes_per_launch": b[dom],
            "other_kernel_us": round((dual_ms if dom == "primal" else primal_ms) * 1e3, 2)}


def cpu_rate(models, its=(64, 192)) -> tuple[float, str]:
    """Seconds per nonzero-iteration of the oracle restatement of the reference's CPU PDLP (one
    core), from two short runs on each model."""
    rates = []
    for md in models:
        _, per_it, _ = port_rate(md, its)
        rates.append(per_it / md.nnz)
    return float(np.mean(rates)), "port"


def cpu_block(value: float, sample: str) -> dict:
    cal = calibration()
    out = {"value": round(value, 1), "unit": "s", "cores": 1, "kind": "port", "sample": sample,
           **host_info()}
    if cal is not None:
        out["reference_over_port"] = round(cal["reference_over_port"], 4)
        out["reference_equivalent"] = round(value * cal["reference_over_port"], 1)
    return out


def golden(name: str) -> dict | None:
    path = ROOT / "tests" / "golden" / f"{name}.json"
    return json.loads(path.read_text()) if path.exists() else None


def c4_result(args, fp64_every: int = 16, cpu_probe_every: int = 4) -> dict:
    """The 256 blending LPs as one stacked batch: the fields of its line (value = seconds from the
    batch resident on the GPU until the last LP stopped). FP64 one LP at a time on every
    `fp64_every`-th LP, scaled; their objectives also against HiGHS where c4_sample.json has it."""
    import samaya_amd as smy
    import samaya_amd.device as D
    import scipy.sparse as sp
    t0 = time.perf_counter()
    K = int(256 * args.scale) if args.scale < 1.0 else 256
    models = [smy.generators.blending(4000 + i) for i in range(K)]
    nms = [md.native() for md in models]
    nnz = sum(md.nnz for md in models)
    log(f"{K} blending LPs, rows {min(md.m for md in models)}..{max(md.m for md in models)}, "
        f"nnz {nnz}, generated in {time.perf_counter() - t0:.1f} s")
    M = sum(md.m for md in models)
    N = sum(md.n for md in models)

    def step_kernels() -> dict:
        """Live timing of the FP32 step kernels over the whole stacked batch."""
        As, Ats, costs, los, ups = [], [], [], [], []
        for md in models:
            A = sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))
            T = A.T.tocsc()
            T.sort_indices()
            As.append((md.m, md.n, md.start, md.index, md.value))
            Ats.append((md.n, md.m, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data))
            costs.append(np.concatenate([md.obj, np.zeros(md.m)]))
            los.append(np.concatenate([md.col_lower, md.row_lower]))
            ups.append(np.concatenate([md.col_upper, md.row_upper]))
        dev = D.PdhgBatch.create(As, Ats, costs, los, ups, "fp32")
        if args.fp64_accumulation:
            dev.set_fp64_accumulation(True)
        out = batch_roofline(dev, nnz, M, N, None, args.adaptive_c4,
                             traffic_of="c4" if args.scale == 1.0 else None,
                             halpern=args.reflection if args.halpern_c4 else None)
        # The tuner's choice on this handle: a tile per size bin, and whether the bins go apart.
        out["size_bins"] = {}
        for matrix, name in ((0, "A"), (1, "At")):
            bins = [dev.bin_tile(matrix, "fp32", b) for b in range(3)]
            out["size_bins"][name] = {"apart": bool(bins[0][2]) if bins[0] else False,
                                      "tiles": [list(t[:2]) if t else None for t in bins],
                                      "one_tile": list(dev.tile(matrix, "fp32") or ())}
        del dev
        return out

    if args.kernels_only:
        return {"kernels_only": True, "roofline": step_kernels(),
                "config": {"workload": f"{K} blending LPs, stacked FP32 step kernels"}}
    opts = smy.PdlpOptions(tol=args.tol, precision="mixed", algorithm=args.algorithm, verify=True,
                           fp32_switch_tol=args.fp32_switch_tol, log_level=1 if args.verbose else 0,
                           adaptive_steps=args.adaptive_c4, fp64_accumulation=args.fp64_accumulation,
                           halpern=args.halpern_c4, reflection=args.reflection)
    smy.solve_lp_batch(nms[:8], opts)  # warm up the tuner and the graph path on a small batch
    runs = []
    t_start = time.perf_counter()
    for k in range(args.sub_steps):
        res, times = smy.solve_lp_batch(nms, opts)
        runs.append((res, times))
        its = np.array([r.iterations for r in res])
        log(f"c4 step {k}: batch {times.solve_seconds - times.setup_seconds:.2f} s after setup "
            f"{times.setup_seconds:.2f} s; iterations median {int(np.median(its))} max {its.max()}; "
            f"statuses {sorted(set(r.status_name for r in res))}; verified {sum(r.verified for r in res)}")
    wall = time.perf_counter() - t_start
    res, times = runs[-1]
    value = float(np.mean([t.solve_seconds - t.setup_seconds for _, t in runs]))
    assert all(r.status == smy.OPTIMAL for r in res), [r.message for r in res if r.status != smy.OPTIMAL][:3]

    fp64 = None
    if not args.no_fp64:
        # Plain FP64 PDLP (the reference's algorithm), one LP at a time: every fp64_every-th LP.
        sample = list(range(0, K, max(1, fp64_every)))
        o64 = smy.PdlpOptions(tol=args.tol, precision="fp64", algorithm="reference", verify=True)
        secs, its64, objs = [], [], []
        for i in sample:
            r = smy.solve_lp(nms[i], o64)
            secs.append(r.solve_seconds - r.setup_seconds)
            its64.append(r.iterations)
            objs.append(abs(res[i].objective - r.objective) / (1 + abs(r.objective)))
        fp64 = {"seconds": sum(secs) * K / len(sample), "sample": sample, "its": its64,
                "obj_err": max(objs)}
        log(f"c4 fp64 one at a time: {sum(secs):.2f} s for {len(sample)} LPs -> {fp64['seconds']:.1f} s")

    rl = step_kernels()
    highs = golden("c4_sample")
    highs_err = None
    if highs is not None and K == 256:
        errs = [abs(res[e["index"]].objective - e["objective"]) / (1 + abs(e["objective"]))
                for e in highs["lps"] if e["status"] == 0]
        highs_err = max(errs) if errs else None

    cpu = None
    if not args.no_cpu_baseline and fp64 is not None:
        probe = [models[i] for i in fp64["sample"][::cpu_probe_every]]
        rate, _ = cpu_rate(probe)
        est = sum(rate * models[i].nnz * it for i, it in zip(fp64["sample"], fp64["its"])) * K / len(fp64["sample"])
        cpu = cpu_block(est, f"oracle restatement of the reference's CPU PDLP, 64 and 192 FP64 "
                             f"iterations on {len(probe)} of the LPs ({rate * 1e9:.2f} ns per "
                             f"nonzero-iteration), times the FP64 iteration counts of the "
                             f"{len(fp64['sample'])} sampled LPs, scaled to {K}; extrapolated")
    its = np.array([r.iterations for r in res])
    out = {
        "value": round(value, 4), "unit": "s", "higher_is_better": False,
        "ms_per_step": round(wall / len(runs) * 1e3, 1), "steps": len(runs),
        "status": "optimal", "iterations_median": int(np.median(its)), "iterations_max": int(its.max()),
        "fp32_iterations_sum": int(sum(r.fp32_iterations for r in res)),
        "fp32_switch_tol": args.fp32_switch_tol, "adaptive_steps": args.adaptive_c4,
        "method": args.c4_method,
        "setup_s": round(times.setup_seconds, 3), "verified": int(sum(r.verified for r in res)),
        "lps": K, "roofline": rl, "cpu_baseline": cpu,
        "config": {"workload": f"{K} multi-period blending LPs (periods U{{4..48}}, 20 units x "
                               f"30 streams), {M} rows, {N} columns, {nnz} nonzeros; stacked "
                               f"batch, mixed precision, rel KKT {args.tol:g} each",
                   "lps": K, "rows": M, "cols": N, "nnz": nnz, "tol": args.tol},
    }
    if highs_err is not None:
        out["rel_objective_error_vs_highs_max"] = highs_err
        out["highs_sample"] = f"{sum(e['status'] == 0 for e in highs['lps'])} LPs (every 32nd)"
    if fp64 is not None:
        out["fp64_one_at_a_time_s"] = round(fp64["seconds"], 2)
        out["fp64_sample"] = f"{len(fp64['sample'])} of {K} LPs solved one at a time, scaled"
        out["speedup_vs_fp64"] = round(fp64["seconds"] / value, 3)
        out["rel_objective_error_vs_fp64_max"] = fp64["obj_err"]
    return out


def c5_result(args, feasible_sample: int = 32, infeasible_sample: int = 16) -> dict:
    """The 1,024 knapsack node relaxations as one shared-matrix batch, warm-started from the root
    LP: the fields of its line (value = seconds until the last node stopped)."""
    import samaya_amd as smy
    import samaya_amd.device as D
    import scipy.sparse as sp
    t0 = time.perf_counter()
    K = int(1024 * args.scale) if args.scale < 1.0 else 1024
    md, lows, ups = smy.generators.c5(K)
    nm = md.native()
    log(f"knapsack m={md.m} n={md.n} nnz={md.nnz}, {K} bound vectors, {time.perf_counter() - t0:.1f} s")

    def step_kernels() -> dict:
        """Live timing of the shared-matrix FP32 step kernels over all K instances."""
        A = sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))
        T = A.T.tocsc()
        T.sort_indices()
        a = (md.m, md.n, md.start, md.index, md.value)
        at = (md.n, md.m, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data)
        cost = np.concatenate([-md.obj, np.zeros(md.m)])
        lo_full = [np.concatenate([lo, md.row_lower]) for lo in lows]
        up_full = [np.concatenate([up, md.row_upper]) for up in ups]
        coded = shared_bounds_coded(lows, ups)
        dev = D.PdhgBatch.create_shared(a, at, cost, lo_full, up_full, "fp32")
        out = batch_roofline(dev, md.nnz, md.m, md.n, K, coded=coded,
                             traffic_of="c5" if args.scale == 1.0 else None,
                             halpern=args.reflection if args.c5_method == "halpern" else None)
        out["bounds_coded"] = coded
        del dev
        return out

    if args.kernels_only:
        return {"kernels_only": True, "roofline": step_kernels(),
                "config": {"workload": f"{K} node relaxations, shared-matrix FP32 step kernels"}}
    opts = smy.PdlpOptions(tol=args.tol, precision="mixed", algorithm=args.algorithm, verify=True,
                           fp32_switch_tol=args.fp32_switch_tol_c5, log_level=1 if args.verbose else 0,
                           adaptive_steps=False, fp64_accumulation=args.fp64_accumulation,
                           halpern=args.c5_method == "halpern", reflection=args.reflection)
    root = smy.solve_lp(nm, opts)
    log(f"c5 root: {root.status_name} {root.iterations} it, omega {root.omega:.3e}")
    start = (root.col_value, root.row_dual)
    opts.start_omega = root.omega
    smy.solve_lp_bounds_batch(nm, lows[:16], ups[:16], opts, start=start)  # warm up
    runs = []
    t_start = time.perf_counter()
    for k in range(args.sub_steps):
        res, times = smy.solve_lp_bounds_batch(nm, lows, ups, opts, start=start)
        runs.append((res, times))
        its = np.array([r.iterations for r in res])
        log(f"c5 step {k}: batch {times.solve_seconds - times.setup_seconds:.2f} s after setup "
            f"{times.setup_seconds:.2f} s; iterations median {int(np.median(its))} max {its.max()}; "
            f"statuses {sorted(set(r.status_name for r in res))}; verified {sum(r.verified for r in res)}")
    wall = time.perf_counter() - t_start
    res, times = runs[-1]
    value = float(np.mean([t.solve_seconds - t.setup_seconds for _, t in runs]))
    A = sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))
    infeasible = [bool(np.any(A @ lo > md.row_upper + 1e-9)) for lo in lows]
    statuses_ok = all(r.status == (smy.INFEASIBLE if inf else smy.OPTIMAL) for r, inf in zip(res, infeasible))

    fp64 = None
    if not args.no_fp64:
        # Plain FP64 PDLP node by node, warm-started alike; a sample stratified by the nodes'
        # classes (feasible nodes take far fewer iterations than the infeasible ones, which end
        # when the Farkas test succeeds), each class scaled by its count.
        feas = [i for i in range(K) if not infeasible[i]]
        infs = [i for i in range(K) if infeasible[i]]
        strata = [(feas, feas[::max(1, len(feas) // feasible_sample)][:feasible_sample]),
                  (infs, infs[::max(1, len(infs) // infeasible_sample)][:infeasible_sample])]
        o64 = smy.PdlpOptions(tol=args.tol, precision="fp64", algorithm="reference", verify=True,
                              start_omega=root.omega)
        total, its64, objs, sample = 0.0, [], [], []
        for members, chosen in strata:
            if not chosen:
                continue
            secs = []
            for i in chosen:
                r = smy.solve_lp_bounds_batch(nm, [lows[i]], [ups[i]], o64, start=start)[0][0]
                secs.append(r.solve_seconds - r.setup_seconds)
                its64.append(r.iterations)
                sample.append(i)
                if r.status == smy.OPTIMAL and res[i].status == smy.OPTIMAL:
                    objs.append(abs(res[i].objective - r.objective) / (1 + abs(r.objective)))
            total += float(np.mean(secs)) * len(members)
            log(f"c5 fp64 one at a time: {len(chosen)} of {len(members)} nodes, mean {np.mean(secs):.3f} s")
        fp64 = {"seconds": total, "sample": sample, "its": its64, "obj_err": max(objs) if objs else None}
        log(f"c5 fp64 one at a time: {fp64['seconds']:.1f} s estimated for {K} nodes")

    rl = step_kernels()

    cpu = None
    if not args.no_cpu_baseline and fp64 is not None:
        node = smy.Model(md.m, md.n, md.start, md.index, md.value, md.obj, lows[0], ups[0],
                         md.row_lower, md.row_upper, md.maximize, md.offset)
        rate, _ = cpu_rate([node])
        est = rate * md.nnz * float(np.mean(fp64["its"])) * K
        cpu = cpu_block(est, f"oracle restatement of the reference's CPU PDLP, 64 and 192 FP64 "
                             f"iterations on one node ({rate * 1e9:.2f} ns per nonzero-iteration), "
                             f"times the mean FP64 iteration count of the sampled nodes, times {K}; "
                             f"extrapolated")
    its = np.array([r.iterations for r in res])
    out = {
        "value": round(value, 4), "unit": "s", "higher_is_better": False,
        "ms_per_step": round(wall / len(runs) * 1e3, 1), "steps": len(runs),
        "status": "expected statuses" if statuses_ok else "unexpected statuses",
        "optimal": int(sum(r.status == smy.OPTIMAL for r in res)),
        "infeasible": int(sum(r.status == smy.INFEASIBLE for r in res)),
        "infeasible_by_construction": int(sum(infeasible)),
        "iterations_median": int(np.median(its)), "iterations_max": int(its.max()),
        "root_iterations": root.iterations, "setup_s": round(times.setup_seconds, 3),
        "fp32_switch_tol": args.fp32_switch_tol_c5, "adaptive_steps": False,
        "method": args.c5_method,
        "fp32_iterations_sum": int(sum(r.fp32_iterations for r in res)),
        "verified": int(sum(r.verified for r in res)), "roofline": rl, "cpu_baseline": cpu,
        "config": {"workload": f"{K} node relaxations (1-5% of columns fixed) of a knapsack LP "
                               f"m={md.m} n={md.n} nnz={md.nnz}; shared-matrix batch (SpMM), "
                               f"mixed precision, warm-started from the root, rel KKT {args.tol:g}",
                   "nodes": K, "m": md.m, "n": md.n, "nnz": md.nnz, "tol": args.tol},
    }
    if fp64 is not None:
        out["fp64_one_at_a_time_s"] = round(fp64["seconds"], 2)
        out["fp64_sample"] = (f"{len(fp64['sample'])} of {K} nodes solved one at a time, warm-started "
                              f"alike: {feasible_sample} feasible and {infeasible_sample} infeasible, "
                              f"each mean scaled by its class count")
        out["speedup_vs_fp64"] = round(fp64["seconds"] / value, 3)
        out["rel_objective_error_vs_fp64_max"] = fp64["obj_err"]
    return out


def batch_line(args, workload: str) -> int:
    """--workload c4 / c5: the batch's own line (or its kernel timings with --kernels-only)."""
    args.sub_steps = args.steps
    res = c4_result(args) if workload == "c4" else c5_result(args)
    if res.get("kernels_only"):
        print(json.dumps({"workload": res["config"]["workload"], "roofline": res["roofline"]}), flush=True)
        return 0
    value, ms = res.pop("value"), res.pop("ms_per_step")
    for k in ("unit", "higher_is_better", "steps"):
        res.pop(k, None)
    line = base_line(args, value, ms, res["config"]["workload"], res)
    line["warmup"] = 1
    print(json.dumps(line), flush=True)
    return 0


def mip_result(args, steps: int = 1) -> dict:
    """A knapsack MIP (m 10, n 60, all columns binary) by branch and bound whose node relaxations
    are bounded in waves of 256 by batched PDLP (smy_mip_solve): the fields of its line (value =
    seconds until the gap closed at 1e-4). Compared with the same search bounding one node per
    batch (a time-capped sample: nodes per second, scaled to the node count) and with HiGHS's
    optimum (mip_knapsack.json)."""
    import samaya_amd as smy
    md = smy.generators.knapsack(7101, 10, 60, 4)
    lp = smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced", halpern=True, verify=False)
    smy.solve_mip(smy.generators.knapsack(7001, 3, 30, 3), np.ones(30, bool), smy.MipOptions(lp=lp, wave=64))
    runs = []
    for _ in range(max(1, steps)):
        r = smy.solve_mip(md, np.ones(md.n, bool), smy.MipOptions(lp=lp, wave=256))
        runs.append(r)
        log(f"mip: {r.status_name}, objective {r.objective}, bound {r.bound:.6f}, {r.nodes} nodes in "
            f"{r.waves} waves, {r.seconds:.2f} s")
    r = runs[-1]
    one = smy.solve_mip(md, np.ones(md.n, bool), smy.MipOptions(lp=lp, wave=1, time_limit=10.0))
    rate1 = one.nodes / one.seconds
    log(f"mip one node per batch: {one.nodes} nodes in {one.seconds:.1f} s ({rate1:.0f} nodes/s)")
    fx = golden("mip_knapsack") or {}
    highs = fx.get("bench", {}).get("objective")
    value = float(np.mean([t.seconds for t in runs]))
    return {
        "value": round(value, 4), "unit": "s", "higher_is_better": False,
        "status": r.status_name, "objective": r.objective, "bound": r.bound, "gap": r.gap,
        "nodes": r.nodes, "waves": r.waves, "nodes_per_s": round(r.nodes / r.seconds, 1),
        "pruned_bound": r.pruned_bound, "pruned_infeasible": r.pruned_infeasible,
        "lp_iterations": r.lp_iterations,
        "one_node_per_batch_nodes_per_s": round(rate1, 1),
        "one_node_per_batch_estimate_s": round(r.nodes / rate1, 2),
        "speedup_vs_one_node_per_batch": round(r.nodes / rate1 / value, 2),
        "highs_objective": highs,
        "rel_objective_error_vs_highs": (abs(r.objective - highs) / (1 + abs(highs))
                                         if highs is not None else None),
        "roofline": None, "cpu_baseline": None,
        "config": {"workload": "knapsack MIP m=10 n=60 (binary), branch and bound with waves of 256 "
                               "node relaxations on batched PDLP (Halpern, mixed), gap 1e-4"},
    }


def mip_line(args) -> int:
    """--workload mip: the MIP sub-result as its own line."""
    res = mip_result(args, args.steps)
    value = res.pop("value")
    for k in ("unit", "higher_is_better"):
        res.pop(k, None)
    line = base_line(args, value, value * 1e3, res["config"]["workload"], res)
    line["warmup"] = 1
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
