"""Writes the golden fixtures of tests/golden/*.json by running the reference's own PDLP path,
compiled from its sources by oracle/Makefile into oracle/_ref/libsamaya_ref.so.

The instances are the reference's own test inputs: its random LP families
(tests/lp_generators.hpp:33, std::mt19937 seeds of tests/test_pdlp.cpp and tests/test_gpu.cpp),
drawn by the reference's generator code itself, so the fixtures hold exactly the models the
reference's tests solve, with the reference's answers (status, iterations, restart schedule,
scaled PDLP point, model-space answer, verifier verdict) and the dense long-double simplex
objective of tests/reference_lp.hpp. The large CPU case (the 2,000 x 5,000 sparse LP, this
library's generator seed 1001) is solved by the reference at tolerance 1e-6 too.

Run in the container that has the reference tree:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import oracle_lib as O  # noqa: E402

FEASIBLE, RANDOM, DEGENERATE, LOOSE, BOXED = range(5)


def f(a):
    return [float(t) if np.isfinite(t) else ("inf" if t > 0 else "-inf") for t in np.asarray(a)]


def ref_case(md, tol, max_iter, with_simplex=True):
    r = O.ref_solve_lp(md, tol, max_iter)
    case = {"model": md.to_json(), "tol": tol, "max_iterations": max_iter,
            "status": r["status"], "iterations": r["iterations"], "objective": r["objective"],
            "verified": r["verified"], "primal_residual": r["primal_residual"],
            "dual_residual": r["dual_residual"], "gap": r["gap"],
            "restarts": [[a, b] for a, b, _, _ in r["restarts"]],
            "pdlp_x": f(r["pdlp_x"]), "pdlp_y": f(r["pdlp_y"]),
            "col_value": f(r["col_value"]), "row_dual": f(r["row_dual"])}
    if r["status"] in (1, 2):
        case["ray"] = f(r["ray"][: md.m if r["status"] == 1 else md.n])
    if with_simplex:
        st, obj = O.ref_reference_lp(md)
        case["simplex_status"], case["simplex_objective"] = st, obj
    return case


def family_cases(seed, family, max_rows, max_cols, count, tol, max_iter, keep=None):
    rng = O.RefRng(seed)
    out = []
    for k in range(count):
        md = rng.random_lp(family, max_rows, max_cols)
        if keep is not None:
            st, obj = O.ref_reference_lp(md)
            if st != keep:
                continue
        c = ref_case(md, tol, max_iter)
        c["k"] = k
        out.append(c)
    return out


def write(name, payload):
    path = HERE / f"{name}.json"
    path.write_text(json.dumps(payload, separators=(",", ":")))
    print(f"wrote {path} ({path.stat().st_size} bytes)")


def main():
    assert O.ref_available(), "build the reference first: make -C oracle ref"
    # tests/test_pdlp.cpp:85-232 (seed 42, 50 LPs <= 8x8, tol 1e-8, 100000 iterations).
    write("pdlp_random_feasible", {
        "source": "tests/test_pdlp.cpp:85-232", "seed": 42,
        "cases": family_cases(42, FEASIBLE, 8, 8, 50, 1e-8, 100000)})
    # tests/test_pdlp.cpp:234-400 (seed 123, 30 LPs <= 10x10, tol 1e-4).
    write("pdlp_vs_simplex", {
        "source": "tests/test_pdlp.cpp:234-400", "seed": 123,
        "cases": family_cases(123, FEASIBLE, 10, 10, 30, 1e-4, 100000)})
    # tests/test_pdlp.cpp:402-478 (seed 4242: 30 kRandom kept when infeasible, then 30 kLoose
    # kept when unbounded, one generator; tol 1e-4, 50000 iterations).
    rng = O.RefRng(4242)
    infeasible, unbounded = [], []
    for k in range(30):
        md = rng.random_lp(RANDOM, 8, 8)
        if O.ref_reference_lp(md)[0] == 1:
            c = ref_case(md, 1e-4, 50000)
            c["k"] = k
            infeasible.append(c)
    for k in range(30):
        md = rng.random_lp(LOOSE, 8, 8)
        if O.ref_reference_lp(md)[0] == 2:
            c = ref_case(md, 1e-4, 50000)
            c["k"] = k
            unbounded.append(c)
    write("pdlp_infeasible_unbounded", {"source": "tests/test_pdlp.cpp:402-478", "seed": 4242,
                                         "infeasible": infeasible, "unbounded": unbounded})
    # tests/test_pdlp.cpp:497-540 (seed 515, 60 LPs <= 8x8, the solver's tolerance 1e-8 and
    # automatic iteration limit).
    write("pdlp_through_solver", {
        "source": "tests/test_pdlp.cpp:497-540", "seed": 515,
        "cases": family_cases(515, FEASIBLE, 8, 8, 60, 1e-8, -1)})
    # tests/test_gpu.cpp:407-477 (seed 42, 20 LPs <= 6x6 kept when optimal, tol 1e-6, 10000).
    write("gpu_vs_cpu", {"source": "tests/test_gpu.cpp:407-477", "seed": 42,
                         "cases": family_cases(42, FEASIBLE, 6, 6, 20, 1e-6, 10000, keep=0)})
    # tests/test_gpu.cpp:514-638 and :640-673: models with random equilibration scales
    # U(0.5, 2) drawn after each model from the same generator (seeds 611 and 612).
    for seed, count, name in ((611, 20, "gpu_evaluate"), (612, 1, "gpu_restart")):
        rng = O.RefRng(seed)
        lib = O.ref()
        lib.ref_uniform.argtypes = [C.c_void_p, C.c_double, C.c_double]
        lib.ref_uniform.restype = C.c_double
        cases = []
        for k in range(count):
            md = rng.random_lp(FEASIBLE, 20, 20)
            rs = [lib.ref_uniform(rng.h, 0.5, 2.0) for _ in range(md.m)]
            cs = [lib.ref_uniform(rng.h, 0.5, 2.0) for _ in range(md.n)]
            cases.append({"model": md.to_json(), "rs": rs, "cs": cs})
        write(name, {"source": "tests/test_gpu.cpp:514-673", "seed": seed, "cases": cases})
    # tests/test_gpu.cpp:675-697 (seed 613, 5 LPs <= 30x30, tol 1e-6, 20000).
    write("gpu_repeatable", {"source": "tests/test_gpu.cpp:675-697", "seed": 613,
                             "cases": family_cases(613, FEASIBLE, 30, 30, 5, 1e-6, 20000)})
    # The 2,000 x 5,000 sparse LP (this library's generator, seed 1001), reference PDLP at 1e-6.
    import samaya_amd as smy
    md = O.M.from_product(smy.generators.c1())
    t0 = time.perf_counter()
    r = O.ref_solve_lp(md, 1e-6, -1)
    seconds = time.perf_counter() - t0
    write("c1_reference", {
        "source": "reference solve_pdlp (CPU) on generators.c1()", "tol": 1e-6,
        "status": r["status"], "iterations": r["iterations"], "objective": r["objective"],
        "verified": r["verified"], "primal_residual": r["primal_residual"],
        "dual_residual": r["dual_residual"], "gap": r["gap"],
        "restarts": [[a, b] for a, b, _, _ in r["restarts"]], "seconds_this_container": seconds,
        "nnz": md.to_json()["start"][-1], "col_value_head": f(r["col_value"][:50]),
        "model_sha": __import__("hashlib").sha256(np.asarray(md.value).tobytes()).hexdigest()})


if __name__ == "__main__":
    main()
