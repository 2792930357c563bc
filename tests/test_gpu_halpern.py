"""Restarted Halpern PDHG with reflection on the GPU (no reference counterpart; Lu and Yang,
"Restarted Halpern PDHG for linear programming", 2024). The step is pinned by hand algebra on a
one-variable LP, by its fixed point, by the plain PDHG operator it reduces to at reflection 0 and
lambda -> 1, and by a NumPy restatement over many steps; the solver by the reference's own test
families (statuses as the reference's PDLP, objectives as the simplex's) and the 2,000 x 5,000 LP
against HiGHS."""
import numpy as np
import pytest
import scipy.sparse as sp

import samaya_amd as smy
import samaya_amd.device as D
from conftest import load_golden
from oracle_lib import LP, M

pytestmark = pytest.mark.gpu
CURRENT, AVERAGE = 0, 1


def device_of(lp: LP, precision="fp64"):
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n))
    T = A.T.tocsc()
    T.sort_indices()
    return D.PdhgDevice.create((lp.m, lp.n, lp.start, lp.index, lp.value),
                               (lp.n, lp.m, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data),
                               lp.cost, lp.lower, lp.upper, precision)


def halpern_reference(lp: LP, v, y, v0, y0, tau, sigma, k, rho):
    """One step of the published iteration, NumPy: (z', T(z))."""
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n))
    n = lp.n
    kty = np.concatenate([A.T @ y, -y])
    vp = np.clip(v - tau * (lp.cost + kty), lp.lower, lp.upper)
    w = 2 * vp - v
    yp = y + sigma * (A @ w[:n] - w[n:])
    lam = (k + 1) / (k + 2)
    return (lam * ((1 + rho) * vp - rho * v) + (1 - lam) * v0,
            lam * ((1 + rho) * yp - rho * y) + (1 - lam) * y0, vp, yp)


def tiny_lp():
    # min 2 x  s.t.  r = 3 x, 1 <= r <= 6, 0 <= x <= 4   (v = (x, r), K = [3 -1])
    return LP(1, 1, np.array([0, 1], np.int64), np.array([0], np.int32), np.array([3.0]),
              np.array([2.0, 0.0]), np.array([0.0, 1.0]), np.array([4.0, 6.0]))


def test_one_step_by_hand(gpu):
    # From z = (x, r, y) = (1, 5, 0.5), tau = 0.1, sigma = 0.2, k = 0 (lambda = 1/2), rho = 1:
    #   x+ = clamp(1 - 0.1 (2 + 3 * 0.5)) = 0.65,   r+ = clamp(5 - 0.1 (0 - 0.5)) = 5.05
    #   w = 2 v+ - v = (0.3, 5.1),   y+ = 0.5 + 0.2 (3 * 0.3 - 5.1) = -0.34
    lp = tiny_lp()
    dev = device_of(lp)
    dev.set_halpern(True, 1.0)
    # set_start puts the start point in the restart point too: the anchor is z itself, so
    # z' = (2 T(z) - z) / 2 + z / 2 = T(z).
    dev.set_start([1.0, 5.0], [0.5])
    dev.step(0.1, 0.2, 1.0, 0)
    s = dev.download()
    np.testing.assert_allclose(s["v_new"], [0.65, 5.05], rtol=0, atol=1e-14)
    np.testing.assert_allclose(s["y_new"], [-0.34], rtol=0, atol=1e-14)
    np.testing.assert_allclose(s["v_erg"], [0.65, 5.05], rtol=0, atol=1e-14)
    np.testing.assert_allclose(s["y_erg"], [-0.34], rtol=0, atol=1e-14)
    dv, dy = dev.fixed_point()
    assert abs(dv - (0.35 ** 2 + 0.05 ** 2)) <= 1e-14 and abs(dy - 0.84 ** 2) <= 1e-14
    # A second step (k = 1, lambda = 2/3) from z' = T(z) with anchor (1, 5, 0.5):
    dev.advance()
    dev.step(0.1, 0.2, 1.0, 1)
    s2 = dev.download()
    v1, y1 = np.array([0.65, 5.05]), np.array([-0.34])
    ref = halpern_reference(lp, v1, y1, np.array([1.0, 5.0]), np.array([0.5]), 0.1, 0.2, 1, 1.0)
    # by hand: x+ = clamp(0.65 - 0.1 (2 - 1.02)) = 0.552, r+ = clamp(5.05 - 0.1 (0.34)) = 5.016
    assert abs(ref[2][0] - 0.552) < 1e-14 and abs(ref[2][1] - 5.016) < 1e-14
    np.testing.assert_allclose(s2["v_new"], ref[0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(s2["y_new"], ref[1], rtol=0, atol=1e-14)


def test_optimal_point_is_fixed(gpu):
    # At a primal-dual optimal point z* the PDHG operator returns z*, so Halpern steps anchored
    # at z* stay there. min -x1 - x2 s.t. x1 + x2 = r <= 1, 0 <= x <= 1: x = (0.5, 0.5), r = 1
    # and y = 1 (K = [A -I]: c + K'y = (0, 0, -1), which pushes r against its upper bound) is
    # optimal.
    lp = LP(1, 2, np.array([0, 1, 2], np.int64), np.array([0, 0], np.int32), np.array([1.0, 1.0]),
            np.array([-1.0, -1.0, 0.0]), np.array([0.0, 0.0, -np.inf]), np.array([1.0, 1.0, 1.0]))
    dev = device_of(lp)
    dev.set_halpern(True, 0.7)
    dev.set_start([0.5, 0.5, 1.0], [1.0])
    dev.run(0.3, 0.4, 1.0, 0, 10)
    s = dev.download()
    np.testing.assert_allclose(s["v_new"], [0.5, 0.5, 1.0], atol=1e-15)
    np.testing.assert_allclose(s["y_new"], [1.0], atol=1e-15)
    dv, dy = dev.fixed_point()
    assert dv <= 1e-28 and dy <= 1e-28


def random_lp(seed, m, n):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=0.2, random_state=seed, format="csc")
    A.data = rng.normal(size=A.nnz)
    cost = np.concatenate([rng.uniform(-1, 1, n), np.zeros(m)])
    lower = np.concatenate([np.zeros(n), np.where(rng.random(m) < 0.5, -np.inf, -1.0)])
    upper = np.concatenate([np.full(n, 5.0), rng.uniform(0, 2, m)])
    return LP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, cost, lower, upper)


@pytest.mark.parametrize("rho", [0.0, 0.5, 1.0])
def test_steps_match_restatement(gpu, rho):
    # 64 steps (one captured run) and a restart, FP64, against the NumPy restatement.
    lp = random_lp(7, 40, 90)
    dev = device_of(lp)
    dev.set_halpern(True, rho)
    rng = np.random.default_rng(8)
    v = np.clip(rng.uniform(-1, 3, lp.n + lp.m), lp.lower, lp.upper)
    y = rng.normal(size=lp.m)
    dev.set_start(v, y)
    v0, y0 = v.copy(), y.copy()
    tau, sigma = 0.05, 0.08
    dev.run(tau, sigma, 1.0, 0, 64)
    vp = yp = None
    for k in range(64):
        vn, yn, vp, yp = halpern_reference(lp, v, y, v0, y0, tau, sigma, k, rho)
        if k < 63:
            v, y = vn, yn
    s = dev.download()
    np.testing.assert_allclose(s["v_new"], vn, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(s["y_new"], yn, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(s["v_erg"], vp, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(s["y_erg"], yp, rtol=1e-11, atol=1e-11)
    dv, dy = dev.fixed_point()
    assert abs(dv - np.sum((vp - v) ** 2)) <= 1e-10 * (1 + dv)
    assert abs(dy - np.sum((yp - y) ** 2)) <= 1e-10 * (1 + dy)
    # ||T(z) - anchor||^2 is the restart distance of the "average" slot.
    ddv, ddy = dev.restart_distance(AVERAGE)
    assert abs(ddv - np.sum((vp - v0) ** 2)) <= 1e-10 * (1 + ddv)
    assert abs(ddy - np.sum((yp - y0) ** 2)) <= 1e-10 * (1 + ddy)
    # A restart moves to T(z), which becomes the anchor; the next step has k = 0.
    dev.restart_to(AVERAGE)
    dev.advance()
    dev.step(tau, sigma, 1.0, 0)
    vn2, yn2, _, _ = halpern_reference(lp, vp, yp, vp, yp, tau, sigma, 0, rho)
    s = dev.download()
    np.testing.assert_allclose(s["v_new"], vn2, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(s["y_new"], yn2, rtol=1e-11, atol=1e-11)


def test_reduces_to_pdhg_operator(gpu):
    # reflection 0 and lambda -> 1 (k large): z' -> T(z), the primal-first PDHG step, whose dual
    # part is the reference's dual update with the extrapolation 2 v+ - v.
    lp = random_lp(9, 30, 70)
    dev = device_of(lp)
    dev.set_halpern(True, 0.0)
    v = np.clip(np.linspace(-1, 4, lp.n + lp.m), lp.lower, lp.upper)
    y = np.linspace(-1, 1, lp.m)
    dev.set_start(v, y)
    k = 10 ** 12
    dev.step(0.05, 0.05, 1.0, k)
    s = dev.download()
    np.testing.assert_allclose(s["v_new"], s["v_erg"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(s["y_new"], s["y_erg"], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_halpern_solver_on_the_reference_families(gpu, precision):
    # The reference's random feasible family (tests/test_pdlp.cpp:85-232) at 1e-8: statuses as the
    # reference's PDLP, objectives within 1e-6 of the simplex's; its infeasible and unbounded
    # models at 1e-4 (tests/test_pdlp.cpp:402-478): never optimal, detections counted.
    opts = smy.PdlpOptions(tol=1e-8, max_iterations=100000, halpern=True, precision=precision)
    optimal = 0
    its = []
    for c in load_golden("pdlp_random_feasible")["cases"]:
        r = smy.solve_lp(M.from_json(c["model"]).product_model(), opts)
        if c["status"] == smy.OPTIMAL:
            assert r.status == smy.OPTIMAL, (c["k"], r.status, r.message)
            assert abs(r.objective - c["simplex_objective"]) <= 1e-6 * (1 + abs(c["simplex_objective"]))
            optimal += 1
            its.append((r.iterations, c["iterations"]))
        else:
            assert r.status != smy.OPTIMAL, (c["k"], r.status)
    assert optimal >= 30
    print(f"halpern {precision}: mean iterations {np.mean([a for a, _ in its]):.0f} "
          f"(reference PDLP {np.mean([b for _, b in its]):.0f})")
    fx = load_golden("pdlp_infeasible_unbounded")
    found = {}
    for key in ("infeasible", "unbounded"):
        hits = 0
        for c in fx[key]:
            r = smy.solve_lp(M.from_json(c["model"]).product_model(),
                             smy.PdlpOptions(tol=1e-4, max_iterations=50000, halpern=True,
                                             precision=precision))
            assert r.status != smy.OPTIMAL, (key, c["k"])
            hits += r.status == c["status"]
        found[key] = (hits, len(fx[key]))
    print(f"halpern {precision}, detected:", found)


def test_halpern_c1_against_highs(gpu):
    ref = load_golden("c1_reference")
    highs = ref["highs_objective"]
    md = smy.generators.c1()
    for precision in ("fp64", "mixed"):
        r = smy.solve_lp(md, smy.PdlpOptions(tol=1e-6, halpern=True, precision=precision))
        print(f"halpern c1 {precision}: {r.iterations} it (reference {ref['iterations']}), "
              f"rel objective error vs HiGHS {abs(r.objective - highs) / (1 + abs(highs)):.2e}")
        assert r.status == smy.OPTIMAL
        assert max(r.primal_residual, r.dual_residual, abs(r.gap)) < 1e-6
        assert abs(r.objective - highs) <= 1e-6 * (1 + abs(highs))


def test_halpern_answer_polished_to_the_verifier(gpu):
    # A Halpern answer rejected by the verifier is polished by FP64 PDHG steps with averages (the
    # reference's iteration) at a tenth of the tolerance: the 2,000 x 5,000 LP then verifies, with
    # HiGHS's objective to 1e-6.
    highs = load_golden("c1_reference")["highs_objective"]
    r = smy.solve_lp(smy.generators.c1(), smy.PdlpOptions(tol=1e-6, halpern=True, precision="mixed",
                                                          polish_rounds=3))
    assert r.status == smy.OPTIMAL and r.verified, r.message
    assert abs(r.objective - highs) <= 1e-6 * (1 + abs(highs))


def test_halpern_batch_equals_single_solves(gpu):
    G = smy.generators
    mods = [G.sparse_uniform(121, 60, 150, 4), G.sparse_powerlaw(122, 300, 700, 40, 5.0),
            G.blending(123, periods=2, units=3, streams=4, qualities=2)]
    opts = smy.PdlpOptions(tol=1e-6, max_iterations=50000, precision="mixed", algorithm="enhanced",
                           halpern=True)
    singles = [smy.solve_lp(md, opts) for md in mods]
    batch, _ = smy.solve_lp_batch(mods, opts)
    for a, b in zip(batch, singles):
        assert a.status == b.status == smy.OPTIMAL
        assert a.iterations == b.iterations
        assert a.objective == b.objective
        np.testing.assert_array_equal(a.col_value, b.col_value)
