"""The PDLP solver on the GPU against the reference's answers (golden fixtures written by the
reference's own PDLP, oracle/_ref) and the oracle, on the reference's test families
(tests/test_pdlp.cpp, tests/test_gpu.cpp) and the 2,000 x 5,000 sparse LP."""
import math

import numpy as np
import pytest

import samaya_amd as smy
from conftest import load_golden
from oracle_lib import M, make_problem, oracle_solve_pdlp

pytestmark = pytest.mark.gpu


def model(case):
    return M.from_json(case["model"])


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


def test_pdlp_basic(gpu):
    # tests/test_pdlp.cpp:54-83: min x, 1 <= x <= 2, no rows -> x = 1 at tol 1e-4.
    lp = smy.LpProblem(0, 1, np.array([0, 0]), np.zeros(0, np.int32), np.zeros(0), [1.0], [1.0], [2.0])
    r = smy.solve_pdlp(lp, smy.PdlpOptions(tol=1e-4, max_iterations=10000))
    assert r.status == smy.OPTIMAL and r.iterations > 0
    assert r.primal_residual < 1e-4 and r.dual_residual < 1e-4 and r.gap < 1e-4
    assert abs(r.x[0] - 1.0) <= 1e-4


def test_random_feasible_matches_reference(gpu):
    # tests/test_pdlp.cpp:85-232: every answer verified, objectives equal to the simplex, mean
    # iterations to 1e-8 within the reference's gate (<= 1500); status and iteration count equal
    # to the reference's own run on the same model.
    fx = load_golden("pdlp_random_feasible")
    its, same_iters = [], 0
    for c in fx["cases"]:
        r = smy.solve_lp(model(c).product_model(), smy.PdlpOptions(tol=1e-8, max_iterations=100000))
        assert r.status == c["status"], (c["k"], r.status, c["status"])
        assert r.verified == c["verified"]
        if c["status"] == smy.OPTIMAL:
            its.append(r.iterations)
            assert rel(r.objective, c["simplex_objective"]) <= 1e-6
            same_iters += r.iterations == c["iterations"]
    assert np.mean(its) <= 1500
    assert same_iters >= 0.8 * len(its)  # rounding-level differences may move a restart


def test_vs_simplex_tol_1e4(gpu):
    # tests/test_pdlp.cpp:234-400: tol 1e-4. The reference counts the optima its verifier accepts
    # at 2e-4 (14 for its own answers: at 1e-4 many reduced costs sit just above 2e-4) and
    # requires no status mismatch; here every status equals the reference PDLP's, the objectives
    # agree to the tolerance, and the verifier accepts as many answers as it accepts the
    # reference's, give or take borderline cases.
    fx = load_golden("pdlp_vs_simplex")
    ours = theirs = 0
    for c in fx["cases"]:
        md = model(c)
        pm = md.product_model()
        r = smy.solve_lp(pm, smy.PdlpOptions(tol=1e-4, max_iterations=100000, verify=False))
        assert r.status == c["status"]
        if r.status == smy.OPTIMAL:
            assert rel(r.objective, c["objective"]) <= 1e-3
            ours += smy.verify_lp_optimality(pm, r.col_value, r.row_dual, 2e-4, 2e-4)[0]
            ref_x = np.array([float(t) for t in c["col_value"]])
            ref_y = np.array([float(t) for t in c["row_dual"]])
            theirs += smy.verify_lp_optimality(pm, ref_x, ref_y, 2e-4, 2e-4)[0]
    assert theirs >= 10 and abs(ours - theirs) <= 3


def test_infeasible_and_unbounded_detected_and_verified(gpu):
    # tests/test_pdlp.cpp:402-478: the same status as the reference's PDLP on every model, and
    # certificates accepted by the verifier about as often as the reference's (it requires at
    # least one; its own Farkas certificates verify on 6 of 19 models at tol 1e-4).
    fx = load_golden("pdlp_infeasible_unbounded")
    ok = {"infeasible": [0, 0], "unbounded": [0, 0]}
    for key, want in (("infeasible", smy.INFEASIBLE), ("unbounded", smy.UNBOUNDED)):
        for c in fx[key]:
            md = model(c).product_model()
            r = smy.solve_lp(md, smy.PdlpOptions(tol=1e-4, max_iterations=50000))
            assert r.status == c["status"], (key, c["k"], r.status, r.message)
            ok[key][0] += r.status == want and r.verified
            ok[key][1] += c["verified"]
    for key in ok:
        ours, theirs = ok[key]
        assert ours >= max(1, theirs - 2), (key, ours, theirs)


def test_through_solver_objectives(gpu):
    # tests/test_pdlp.cpp:497-540: objectives within 1e-6 of the simplex; >= 80% of the optima
    # verified from PDLP alone.
    fx = load_golden("pdlp_through_solver")
    optimal = by_pdlp = 0
    for c in fx["cases"]:
        r = smy.solve_lp(model(c).product_model(), smy.PdlpOptions(tol=1e-8))
        if c["simplex_status"] == 0:
            optimal += 1
            if r.status == smy.OPTIMAL and r.verified:
                by_pdlp += 1
                assert rel(r.objective, c["simplex_objective"]) <= 1e-6
        elif c["simplex_status"] == 1:
            assert r.status != smy.OPTIMAL
    assert by_pdlp * 10 >= optimal * 8


def test_gpu_matches_cpu_solution(gpu):
    # tests/test_gpu.cpp:407-477: x within 1e-4 of the CPU PDLP at tol 1e-6.
    fx = load_golden("gpu_vs_cpu")
    for c in fx["cases"]:
        md = model(c)
        lp, _, _ = make_problem(md)
        r = smy.solve_pdlp(lp.product(), smy.PdlpOptions(tol=1e-6, max_iterations=10000))
        cpu = np.array([float(t) for t in c["pdlp_x"]])
        assert r.status == c["status"]
        if r.status == smy.OPTIMAL:
            scale = np.maximum(1.0, np.maximum(np.abs(cpu), np.abs(r.x)))
            assert np.all(np.abs(r.x[:md.n] - cpu[:md.n]) / scale[:md.n] <= 1e-4)


def test_runs_are_repeatable(gpu):
    # tests/test_gpu.cpp:675-697: equal iterations, bit-identical x and y.
    fx = load_golden("gpu_repeatable")
    for c in fx["cases"]:
        lp, _, _ = make_problem(model(c))
        o = smy.PdlpOptions(tol=1e-6, max_iterations=20000)
        a = smy.solve_pdlp(lp.product(), o)
        b = smy.solve_pdlp(lp.product(), o)
        assert a.status == b.status and a.iterations == b.iterations
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_graph_and_single_step_runs_agree(gpu):
    fx = load_golden("gpu_repeatable")
    lp, _, _ = make_problem(model(fx["cases"][0]))
    a = smy.solve_pdlp(lp.product(), smy.PdlpOptions(tol=1e-6, max_iterations=20000, use_graph=True))
    b = smy.solve_pdlp(lp.product(), smy.PdlpOptions(tol=1e-6, max_iterations=20000, use_graph=False))
    assert a.iterations == b.iterations and np.array_equal(a.x, b.x)


def test_decisions_track_the_oracle(gpu):
    # Same restart schedule as the oracle (CPU restatement pinned to the reference) on most
    # models; the device's sums differ from the host's only in rounding.
    fx = load_golden("pdlp_random_feasible")
    same = total = 0
    for c in fx["cases"][:20]:
        lp, _, _ = make_problem(model(c))
        g = smy.solve_pdlp(lp.product(), smy.PdlpOptions(tol=1e-8, max_iterations=100000))
        o = oracle_solve_pdlp(lp, 1e-8, 100000)
        total += 1
        same += [(a, b) for a, b, *_ in g.restarts] == [(a, b) for a, b, *_ in o.restarts]
        assert g.status == o.status
        assert math.isclose(g.operator_norm, o.L, rel_tol=1e-12)
    assert same >= 0.8 * total


def test_c1_sparse_lp_to_1e6(gpu):
    # The 2,000 x 5,000 sparse LP: after the polish the verifier accepts the answer and its
    # objective is within 1e-6 of HiGHS's (golden c1_reference, highs-ipm); the plain 1e-6 answer
    # is within 1e-6 too, as is the reference's own CPU PDLP at 1e-6 (same fixture).
    ref = load_golden("c1_reference")
    highs = ref["highs_objective"]
    md = smy.generators.c1()
    r = smy.solve_lp(md, smy.PdlpOptions(tol=1e-6, polish_rounds=3))
    assert r.status == smy.OPTIMAL
    assert r.verified, r.message
    assert abs(r.objective - highs) <= 1e-6 * (1 + abs(highs))
    plain = smy.solve_lp(md, smy.PdlpOptions(tol=1e-6))
    assert plain.status == smy.OPTIMAL and plain.iterations > 0
    assert abs(plain.objective - highs) <= 1e-6 * (1 + abs(highs))
    assert abs(ref["objective"] - highs) <= 1e-6 * (1 + abs(highs))


def test_c1_mixed_precision(gpu):
    ref = load_golden("c1_reference")
    highs = ref["highs_objective"]
    md = smy.generators.c1()
    r = smy.solve_lp(md, smy.PdlpOptions(tol=1e-6, precision="mixed", polish_rounds=3))
    assert r.status == smy.OPTIMAL and r.verified, r.message
    assert r.fp32_iterations > 0
    assert abs(r.objective - highs) <= 1e-6 * (1 + abs(highs))
