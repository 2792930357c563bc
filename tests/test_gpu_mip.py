"""Branch and bound over batched PDLP node relaxations (smy_mip_solve) against HiGHS's MILP
(golden mip_knapsack.json, scipy.optimize.milp): the same optimal objective (the data are
integers, so a gap closed at 1e-4 leaves no other value), an incumbent that is integral and
satisfies every row and bound, a bound that brackets it within the gap, and a MIP whose LP is
feasible but which has no integer point reported infeasible."""
import numpy as np
import pytest
import scipy.sparse as sp

import samaya_amd as smy
from conftest import load_golden

pytestmark = pytest.mark.gpu
G = smy.generators


def knapsack_case(case):
    seed, m, n, per_col = case["knapsack"]
    md = G.knapsack(seed, m, n, per_col)
    up = np.full(md.n, float(case["col_upper"]))
    return smy.Model(md.m, md.n, md.start, md.index, md.value, md.obj, md.col_lower, up,
                     md.row_lower, md.row_upper, md.maximize, md.offset)


@pytest.mark.parametrize("halpern", [False, True], ids=["pdlp", "halpern"])
def test_knapsack_mips_match_highs(gpu, halpern):
    fx = load_golden("mip_knapsack")
    waves = 0
    for case in fx["cases"]:
        if "knapsack" not in case:
            continue
        md = knapsack_case(case)
        opts = smy.MipOptions(lp=smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced",
                                                 halpern=halpern, verify=False),
                              gap_tol=1e-4, wave=64)
        r = smy.solve_mip(md, np.ones(md.n, bool), opts)
        want = case["objective"]
        assert r.status == smy.OPTIMAL, (case["knapsack"], r)
        assert abs(r.objective - want) <= 1e-6 * (1 + abs(want)), (case["knapsack"], r.objective, want)
        # A maximization: the bound lies above the incumbent, within the gap.
        assert r.objective - 1e-9 <= r.bound <= r.objective + 1e-4 * (1 + abs(r.objective)) + 1e-9
        x = r.col_value
        assert np.array_equal(x, np.round(x))
        assert np.all(x >= md.col_lower) and np.all(x <= md.col_upper)
        A = sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))
        assert np.all(A @ x <= md.row_upper + 1e-9) and np.all(A @ x >= md.row_lower - 1e-9)
        assert abs(md.obj @ x + md.offset - r.objective) <= 1e-9 * (1 + abs(r.objective))
        assert r.nodes >= 1 and r.unresolved == 0
        waves += r.waves
        print(f"{case['knapsack']} ub {case['col_upper']}: {r.objective:.1f} (HiGHS {want:.1f}), "
              f"{r.nodes} nodes in {r.waves} waves, {r.pruned_bound} pruned by bound, "
              f"{r.pruned_infeasible} infeasible, {r.seconds:.2f} s")
    assert waves > 0  # nodes were bounded in batches


def test_integer_infeasible_mip(gpu):
    # min x1 + x2 s.t. 2 x1 + 2 x2 = 1, x1, x2 integer in [0, 1]: the LP optimum is 0.5, no integer
    # point exists, and every branch ends in an infeasible relaxation.
    fx = load_golden("mip_knapsack")
    case = [c for c in fx["cases"] if "handmade" in c][0]
    assert case["status"] == smy.INFEASIBLE
    md = smy.Model(1, 2, np.array([0, 1, 2], np.int64), np.array([0, 0], np.int32), np.array([2.0, 2.0]),
                   np.array([1.0, 1.0]), np.zeros(2), np.ones(2), np.array([1.0]), np.array([1.0]),
                   False, 0.0)
    r = smy.solve_mip(md, [True, True], smy.MipOptions(lp=smy.PdlpOptions(tol=1e-8, verify=False)))
    assert r.status == smy.INFEASIBLE, r
    assert r.col_value is None and r.pruned_infeasible >= 2


def test_mip_refuses_a_model_without_integer_columns(gpu):
    md = G.knapsack(7001, 3, 30, 3)
    with pytest.raises(smy.NativeError):
        smy.solve_mip(md, np.zeros(md.n, bool))
