"""The benchmark configurations at their full sizes on the GPU (the model for these tests is the
reference's tests/test_gpu.cpp:407-477, which runs its GPU PDLP on its largest generated LPs):

  C2  the 1M-column LP (m 500,000, n 1,000,000, nnz ~12M), plain FP64 PDLP with the reference's
      algorithm, against the reference's own CPU PDLP on the same LP (golden c2_reference);
  C3  the same LP in mixed precision: FP32 restarted Halpern steps with FP64 evaluation, restarts
      and termination (the benchmark's line), and PDLP with adaptive steps;
  C4  256 blending LPs as one stacked batch, against HiGHS on every 32nd LP (golden c4_sample)
      and against the same LPs solved alone;
  C5  1,024 knapsack node relaxations as one shared-matrix batch, warm-started from the root:
      statuses against the nodes' infeasibility worked out exactly on the host, objectives
      against plain FP64 PDLP solving sampled nodes one at a time, and weak-duality bounds.

Every single-LP answer's termination measures are recomputed here from the returned point, in
numpy, so a test does not rest on the solver's own report of them."""
import numpy as np
import pytest
import scipy.sparse as sp

import samaya_amd as smy
from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


def csc(md):
    return sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))


def problem_measures(md, r):
    """The termination measures of a model-space answer, recomputed here in numpy: the answer is
    mapped into make_problem's computational form (power-of-two scaling, src/lp/lp_solver.cpp:28-51,
    exact to undo) and measured as src/lp/pdlp.cpp:411-519 defines it:
      rel primal  ||A x - r|| / (1 + max column 2-norm of A * ||x|| + ||r||)
      rel dual    ||g - lambda|| / (1 + max row 2-norm of A * ||y|| + ||c||),
                  g = c + [A -I]'y, lambda the part the bounds absorb
      rel gap     (c'v - sum lambda * bound) / (1 + |c'v| + |sum lambda * bound|)
    and the largest bound violation of v (relative to 1 + |bound|)."""
    lp = smy.make_problem(md)
    cs, rs = smy.compute_scaling(md)
    n, m = md.n, md.m
    sense = -1.0 if md.maximize else 1.0
    x = r.col_value / cs
    rr = r.row_activity * rs
    y = -sense * r.row_dual / rs
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(m, n))
    v = np.concatenate([x, rr])
    a_colmax = np.sqrt(np.asarray(A.multiply(A).sum(axis=0)).ravel()).max()
    at_rowmax = np.sqrt(np.asarray(A.multiply(A).sum(axis=1)).ravel()).max()
    primal = np.linalg.norm(A @ x - rr) / (1 + a_colmax * np.linalg.norm(x) + np.linalg.norm(rr))
    g = lp.cost + np.concatenate([A.T @ y, -y])
    lam = np.where((g > 0) & np.isfinite(lp.lower), g, np.where((g < 0) & np.isfinite(lp.upper), g, 0.0))
    dual = np.linalg.norm(g - lam) / (1 + at_rowmax * np.linalg.norm(y) + np.linalg.norm(lp.cost))
    pobj = lp.cost @ v
    dobj = np.sum(np.where(lam > 0, lam * np.where(np.isfinite(lp.lower), lp.lower, 0.0), 0.0)) + \
        np.sum(np.where(lam < 0, lam * np.where(np.isfinite(lp.upper), lp.upper, 0.0), 0.0))
    gap = (pobj - dobj) / (1 + abs(pobj) + abs(dobj))
    viol = np.maximum(lp.lower - v, v - lp.upper) / (1 + np.where(np.isfinite(lp.lower), abs(lp.lower), 0)
                                                       + np.where(np.isfinite(lp.upper), abs(lp.upper), 0))
    return primal, dual, gap, max(float(viol.max(initial=0.0)), 0.0)


def assert_within_tolerance(md, r, tol=1e-6):
    """The solver's report and the recomputation agree, and both are inside the tolerance (the
    recomputation sums in another order: a relative slack of 1e-9 on the bound). An answer whose
    last steps ran in FP32 lies within the FP32 roundings of the scaled bounds (a relative 2^-23
    on either side: two FP32 ulps); an FP64 one within the FP64 roundings."""
    primal, dual, gap, bounds = problem_measures(md, r)
    assert max(r.primal_residual, r.dual_residual, abs(r.gap)) < tol
    assert max(primal, dual, abs(gap)) < tol * (1 + 1e-9), (primal, dual, gap)
    assert abs(primal - r.primal_residual) <= 1e-9 * tol + 1e-6 * r.primal_residual
    assert abs(dual - r.dual_residual) <= 1e-9 * tol + 1e-6 * r.dual_residual
    fp32_last = r.fp32_iterations == r.iterations
    assert bounds <= (2.0 ** -22 if fp32_last else 1e-12), bounds
    return primal, dual, gap


@pytest.fixture(scope="module")
def c2_model():
    return smy.generators.c2()


C2_FIXTURE = (GOLDEN / "c2_reference.json").exists()


@pytest.mark.skipif(not C2_FIXTURE, reason="tests/golden/c2_reference.json not written yet "
                                           "(make_golden.py c2: the reference's CPU PDLP, ~1 hour)")
def test_c2_fp64_reference_algorithm(gpu, c2_model):
    # Plain FP64 PDLP (the reference's preconditioner, constant steps, restart rule and
    # detection) on the GPU, against the reference's own CPU PDLP on the same LP at 1e-6.
    ref = load_golden("c2_reference")
    md = c2_model
    assert (md.m, md.n, md.nnz) == (ref["m"], ref["n"], ref["nnz"])
    r = smy.solve_lp(md, smy.PdlpOptions(tol=1e-6, precision="fp64", algorithm="reference"))
    print(f"C2 fp64: {r.status_name} {r.iterations} it (reference {ref['iterations']}), objective "
          f"{r.objective!r} (reference {ref['objective']!r}, rel {rel(r.objective, ref['objective']):.2e}), "
          f"measures {r.primal_residual:.2e} {r.dual_residual:.2e} {r.gap:.2e}, dual bound {r.dual_bound!r}")
    assert r.status == smy.OPTIMAL == ref["status"]
    assert_within_tolerance(md, r)
    # The same algorithm from the same start. The GPU sums in another order, so in general the runs
    # part at the rounding level; on this LP they keep the reference's restart schedule and stop at
    # its iteration (100,352, measured), with objectives equal to ~1e-16.
    assert r.iterations == ref["iterations"]
    assert rel(r.objective, ref["objective"]) <= 1e-9
    # Weak duality: the bound from the answer's duals lies below the objective (minimization).
    assert r.dual_bound <= r.objective + 1e-9 * abs(r.objective)
    assert rel(r.dual_bound, r.objective) <= 1e-4


C3_LINES = {
    # The benchmark's line: restarted Halpern steps in FP32 (values, iterates, sums) to the
    # tolerance, FP64 evaluation, restarts and termination, the reference's preconditioner and the
    # columns by count.
    "halpern": dict(halpern=True, algorithm="reference", column_order="touch",
                    fp32_switch_tol=1e-6, fp32_stall_level=0.0),
    # PDLP with adaptive steps and the enhanced preconditioner.
    "adaptive": dict(adaptive_steps=True, algorithm="enhanced"),
}


@pytest.mark.parametrize("line", sorted(C3_LINES))
def test_c3_mixed_precision(gpu, c2_model, line):
    ref = load_golden("c2_reference") if C2_FIXTURE else None
    md = c2_model
    r = smy.solve_lp(md, smy.PdlpOptions(tol=1e-6, precision="mixed", detect=False, **C3_LINES[line]))
    against = f"rel {rel(r.objective, ref['objective']):.2e} to the reference's FP64" if ref else "no fixture"
    print(f"C3 mixed {line}: {r.status_name} {r.iterations} it ({r.fp32_iterations} FP32), objective "
          f"{r.objective!r} ({against}), measures {r.primal_residual:.2e} {r.dual_residual:.2e} "
          f"{r.gap:.2e}, dual bound {r.dual_bound!r}, verified {r.verified} "
          f"({r.max_primal_violation:.1e}, {r.max_dual_violation:.1e})")
    assert r.status == smy.OPTIMAL
    assert r.fp32_iterations >= 0.9 * r.iterations
    assert_within_tolerance(md, r)
    # What rel KKT 1e-6 allows on this LP. Its normalizers are large (||b|| ~ 2e5, the largest
    # column 2-norm ~ 1e3): the allowed primal residual moves c'x by up to |y|'|Ax - r| and the
    # allowed dual residual moves the weak-duality bound by up to sum_j |d_j - lambda_j| |x_j|,
    # each ~1e-4 of the objective here. Measured: the Halpern line's answer at 1e-6 lies 1.7e-4
    # above its own dual bound, the adaptive line's 1e-5; a Halpern run to 1e-7 brackets the
    # optimum in [-1506808.32, -1506806.35] (its dual bound and objective), 4.7e-5 and 2.7e-5
    # from the two 1e-6 answers. So the answers are checked at 1e-4 against the reference's own
    # FP64 PDLP answer at 1e-6 and at 3e-4 against their rigorous dual bound.
    if ref is not None:
        assert rel(r.objective, ref["objective"]) <= 1e-4
    assert r.dual_bound <= r.objective + 1e-9 * abs(r.objective)
    assert rel(r.dual_bound, r.objective) <= 3e-4


def test_c4_blending_batch(gpu):
    # 256 blending LPs (2.6k-60k rows each) as one stacked batch, mixed precision with restarted
    # Halpern steps and the enhanced preconditioner (the benchmark's batch line): every LP optimal
    # within the tolerance, objectives against HiGHS on every 32nd LP, and sampled members equal
    # to the same LP solved alone (same iterations, bit-equal objective).
    highs = load_golden("c4_sample")
    models = smy.generators.c4()
    nms = [md.native() for md in models]
    opts = smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced", halpern=True)
    res, times = smy.solve_lp_batch(nms, opts)
    its = [r.iterations for r in res]
    print(f"C4: {times.solve_seconds - times.setup_seconds:.2f} s, iterations median "
          f"{int(np.median(its))} max {max(its)}, verified {sum(r.verified for r in res)}")
    assert all(r.status == smy.OPTIMAL for r in res), [r.message for r in res if r.status != smy.OPTIMAL][:3]
    assert all(max(r.primal_residual, r.dual_residual, abs(r.gap)) < 1e-6 for r in res)
    errs = []
    for e in highs["lps"]:
        assert e["status"] == 0
        errs.append(rel(res[e["index"]].objective, e["objective"]))
    print(f"C4 vs HiGHS on {len(errs)} LPs: max rel objective error {max(errs):.2e}")
    assert max(errs) <= 2e-6
    for i in (0, 97, 255):
        alone = smy.solve_lp(nms[i], opts)
        assert alone.iterations == res[i].iterations, i
        assert alone.objective == res[i].objective, i


@pytest.mark.parametrize("halpern", [False, True], ids=["pdlp", "halpern"])
def test_c5_node_batch(gpu, halpern):
    # 1,024 node relaxations of the 20,000 x 50,000 knapsack LP, warm-started from the root LP's
    # answer and primal weight, as one shared-matrix batch (PDLP's steps or Halpern steps).
    md, lows, ups = smy.generators.c5(1024)
    nm = md.native()
    A = csc(md)
    # A node is infeasible exactly when the columns it fixes at 1 already overfill a row: every
    # weight is positive and every free column can sit at 0.
    infeasible = [bool(np.any(A @ lo > md.row_upper)) for lo in lows]
    opts = smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced", fp32_switch_tol=1e-6,
                           halpern=halpern)
    root = smy.solve_lp(nm, opts)
    assert root.status == smy.OPTIMAL
    opts.start_omega = root.omega
    start = (root.col_value, root.row_dual)
    res, times = smy.solve_lp_bounds_batch(nm, lows, ups, opts, start=start)
    n_inf = sum(infeasible)
    print(f"C5: {times.solve_seconds - times.setup_seconds:.2f} s, {n_inf} infeasible by "
          f"construction, statuses {sorted(set(r.status_name for r in res))}, verified "
          f"{sum(r.verified for r in res)} (infeasible {sum(r.verified for r, f in zip(res, infeasible) if f)})")
    assert 0 < n_inf < len(lows)
    for k, (r, inf) in enumerate(zip(res, infeasible)):
        assert r.status == (smy.INFEASIBLE if inf else smy.OPTIMAL), (k, r.status_name, r.message)
    # Every infeasible node's Farkas certificate passes the verifier.
    assert all(r.verified for r, inf in zip(res, infeasible) if inf)
    # Optimal nodes: within the tolerance, and the weak-duality bound (an upper bound here, the
    # model maximizes) from each node's duals lies above its objective and close to it.
    for k, r in enumerate(res):
        if infeasible[k]:
            continue
        assert max(r.primal_residual, r.dual_residual, abs(r.gap)) < 1e-6, k
        assert r.dual_bound >= r.objective - 1e-9 * abs(r.objective), k
        assert rel(r.dual_bound, r.objective) <= 1e-4, (k, r.dual_bound, r.objective)
    # Against plain FP64 PDLP (the reference's algorithm) solving sampled feasible nodes alone,
    # warm-started alike.
    o64 = smy.PdlpOptions(tol=1e-6, precision="fp64", algorithm="reference", start_omega=root.omega)
    feas = [k for k in range(len(lows)) if not infeasible[k]]
    errs = []
    for k in feas[:: len(feas) // 8][:8]:
        r64 = smy.solve_lp_bounds_batch(nm, [lows[k]], [ups[k]], o64, start=start)[0][0]
        assert r64.status == smy.OPTIMAL
        errs.append(rel(res[k].objective, r64.objective))
    print(f"C5 vs FP64 one at a time on {len(errs)} nodes: max rel objective error {max(errs):.2e}")
    assert max(errs) <= 1e-6
