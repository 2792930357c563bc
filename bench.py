#!/usr/bin/env python3
"""Benchmark of the PDLP hot path on one MI355X.

Workload: the 1,000,000-column sparse LP (m 500,000, nnz ~12M, power-law column counts on
[2, 5,000] with mean 12, values N(0,1) x 10^U[-2,2], feasible by construction; the library's
generator, seed 1002), solved to max(rel primal, rel dual, |rel gap|) < 1e-6 with
mixed-precision PDLP (FP32 values / iterates / accumulation in the steps, FP64 evaluation,
restarts and termination, FP64 polish). One step = one such solve.

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

Usage: python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--scale 1.0] [--no-fp64]
       [--no-cpu-baseline] [--no-spmv]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBPS = 8000.0
METRIC = ("time-to-1e-6 rel KKT + SpMV HBM GB/s; FP32/batched vs FP64 PDLP, 1 GPU")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(nnz: int, rows: int, cols: int, s: int) -> int:
    return nnz * (s + 4) + 4 * (rows + 1) + s * cols + s * rows


def step_kernel_bytes(nnz: int, m: int, n: int, s: int) -> dict:
    """Algorithmic HBM bytes of the two fused step kernels (each value / index / offset once,
    the gathered vector once, the per-row and per-column update vectors once)."""
    dual = nnz * (s + 4) + 4 * (m + 1) + s * n + 13 * s * m
    primal = nnz * (s + 4) + 4 * (n + 1) + s * m + 8 * s * n
    return {"dual": dual, "primal": primal}


def profiled_traffic(op: str) -> int | None:
    """Calibrated HBM bytes per launch of a step kernel, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/step_kernels_c2_mixed.md), if present."""
    path = ROOT / "profiles" / "step_kernels_c2_mixed.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == op:
            return int(parts[2])
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


def kernel_roofline(md, precision: str) -> dict:
    """Live per-kernel timing of the two step kernels on the workload's matrix."""
    import numpy as np
    import scipy.sparse as sp

    from samaya_amd.device import PdhgDevice

    A = sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))
    T = A.T.tocsc()
    T.sort_indices()
    nt = md.n + md.m
    lower = np.concatenate([md.col_lower, md.row_lower])
    upper = np.concatenate([md.col_upper, md.row_upper])
    cost = np.concatenate([md.obj, np.zeros(md.m)])
    dev = PdhgDevice.create((md.m, md.n, md.start, md.index, md.value),
                            (md.n, md.m, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data),
                            cost, lower, upper, "fp32" if precision == "fp32" else "fp64")
    dual_ms, primal_ms = dev.time_kernels(50)
    s = 4 if precision == "fp32" else 8
    b = step_kernel_bytes(md.nnz, md.m, md.n, s)
    return {"dual_ms": dual_ms, "primal_ms": primal_ms, "bytes": b}


def cpu_baseline(md, gpu_iterations: int, sample_iters=(64, 128)) -> dict | None:
    """The reference's CPU PDLP (oracle/_ref) on the same LP for two short runs; the difference
    gives the time per iteration, extrapolated to the GPU FP64 run's iteration count."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    kind = "reference" if O.ref_available() else "port"
    m = O.M.from_product(md)
    times = []
    for its in sample_iters:
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_solve_lp(m, 1e-6, its, capture_log=False)
        else:
            O.oracle_solve_lp(m, 1e-6, its)
        times.append(time.perf_counter() - t0)
    per_it = (times[1] - times[0]) / (sample_iters[1] - sample_iters[0])
    setup = times[0] - per_it * sample_iters[0]
    value = setup + per_it * gpu_iterations
    return {"value": round(value, 2), "unit": "s", "cores": 1, "kind": kind,
            "sample": (f"{sample_iters[0]} and {sample_iters[1]} FP64 PDLP iterations of the same "
                       f"LP on one core ({times[0]:.1f} s, {times[1]:.1f} s): "
                       f"{per_it * 1e3:.1f} ms/iteration, setup {setup:.1f} s, extrapolated to "
                       f"{gpu_iterations} iterations")}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-spmv", action="store_true")
    ap.add_argument("--algorithm", default="enhanced", choices=["reference", "enhanced"])
    ap.add_argument("--max-iterations", type=int, default=-1,
                    help="cap each solve (profiling runs); the line then reports no speedup")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the batched and "
              "single-LP shardings are described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        return 2

    import torch

    import samaya_amd as smy

    assert torch.cuda.is_available() and smy.gpu_available(), "no usable MI355X"
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    md = smy.generators.c2(args.scale)
    log(f"LP m={md.m} n={md.n} nnz={md.nnz} generated in {time.perf_counter() - t0:.1f} s")
    nm = md.native()
    opts = smy.PdlpOptions(tol=args.tol, precision="mixed", algorithm=args.algorithm, detect=False,
                           verify=True, polish_rounds=0, max_iterations=args.max_iterations)

    for w in range(args.warmup):
        r = smy.solve_lp(nm, opts)
        log(f"warmup {w}: {r.status_name} {r.iterations} it in {r.solve_seconds:.2f} s")

    torch.cuda.synchronize()
    t_start = time.perf_counter()
    runs = []
    for k in range(args.steps):
        r = smy.solve_lp(nm, opts)
        runs.append(r)
        log(f"step {k}: {r.status_name} {r.iterations} it ({r.fp32_iterations} FP32), "
            f"setup {r.setup_seconds:.2f} s, iterate {r.solve_seconds - r.setup_seconds:.3f} s, "
            f"verified {r.verified} ({r.max_primal_violation:.1e}, {r.max_dual_violation:.1e})")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    iterate_s = sum(r.solve_seconds - r.setup_seconds for r in runs) / len(runs)
    setup_s = sum(r.setup_seconds for r in runs) / len(runs)
    last = runs[-1]
    capped = args.max_iterations > 0
    assert capped or all(r.status == smy.OPTIMAL for r in runs), [r.message for r in runs]

    fp64 = None
    if not args.no_fp64 and not capped:
        # Plain FP64 PDLP: the reference's algorithm and precision on the same GPU.
        f = smy.solve_lp(nm, smy.PdlpOptions(tol=args.tol, precision="fp64", algorithm="reference",
                                             detect=False))
        fp64 = f
        log(f"fp64: {f.status_name} {f.iterations} it, iterate {f.solve_seconds - f.setup_seconds:.3f} s")

    rl = kernel_roofline(md, "fp32")
    dom = "primal" if rl["primal_ms"] >= rl["dual_ms"] else "dual"
    dom_ms = rl[f"{dom}_ms"]
    achieved = rl["bytes"][dom] / (dom_ms * 1e-3) / 1e9
    log(f"step kernels fp32: dual {rl['dual_ms'] * 1e3:.1f} us, primal {rl['primal_ms'] * 1e3:.1f} us")

    spmv = None if args.no_spmv else spmv_sweep(md)
    cpu = None
    if not args.no_cpu_baseline:
        gpu_its = fp64.iterations if fp64 is not None else last.iterations
        cpu = cpu_baseline(md, gpu_its)

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
        "algorithm": args.algorithm,
        "status": last.status_name,
        "data": "synthetic (seeded generator)",
        "config": {"workload": "sparse LP m=500000 n=1000000 nnz~12M power-law columns, "
                               "mixed-precision PDLP to rel KKT 1e-6" if args.scale == 1.0 else
                               f"sparse LP scale {args.scale} (m={md.m}, n={md.n})",
                   "m": md.m, "n": md.n, "nnz": md.nnz, "tol": args.tol},
        "iterations": last.iterations,
        "fp32_iterations": last.fp32_iterations,
        "setup_s": round(setup_s, 3),
        "verified": last.verified,
        "objective": last.objective,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": (profiled_traffic("PrimalColsOp<float>" if dom == "primal" else "DualRowsOp<float>")
                                 if args.scale == 1.0 else None),
                     "kernel": f"csr_rows_kernel<float, {'PrimalColsOp' if dom == 'primal' else 'DualRowsOp'}<float>>",
                     "avg_us": round(dom_ms * 1e3, 2), "bytes_per_launch": rl["bytes"][dom]},
        "cpu_baseline": cpu,
    }
    if fp64 is not None:
        line["fp64_time_to_tol_s"] = round(fp64.solve_seconds - fp64.setup_seconds, 4)
        line["fp64_iterations"] = fp64.iterations
        line["fp64_objective"] = fp64.objective
        line["speedup_vs_fp64"] = round(line["fp64_time_to_tol_s"] / iterate_s, 3)
        line["rel_objective_error_vs_fp64"] = abs(last.objective - fp64.objective) / (1 + abs(fp64.objective))
    if spmv is not None:
        line["spmv"] = spmv
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
