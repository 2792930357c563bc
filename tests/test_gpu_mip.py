"""Branch and bound over batched PDLP node relaxations (smy_mip_solve) against HiGHS's MILP
(golden mip_knapsack.json, scipy.optimize.milp): the same optimal objective (the data are
integers, so a gap closed at 1e-4 leaves no other value), an incumbent that is integral and
satisfies every row and bound, a bound that brackets it within the gap, and MIPs whose LP is
feasible but which have no integer point reported infeasible (by the rows' implied bounds, or by
branching)."""
import numpy as np
import pytest
import scipy.sparse as sp

import samaya_amd as smy
from conftest import load_golden

pytestmark = pytest.mark.gpu
G = smy.generators
GOLDEN_MPS = __import__("pathlib").Path(__file__).parent / "golden" / "mps"


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


def test_the_reference_mip_instance(gpu):
    # mrpl_crude_small.mps from the reference's tests/instances (164 rows, 251 columns, 83 integer,
    # columns not all boxed): the rows' implied bounds box every column, so the safe dual bounds
    # prune, and the search closes its gap around HiGHS's optimum: the incumbent (a feasible point)
    # no better than it, the bound (rigorous) no worse, the two within the gap tolerance.
    fx = load_golden("mip_knapsack")
    want = [c for c in fx["cases"] if c.get("mps") == "mrpl_crude_small.mps"][0]["objective"]
    md = smy.read_mps_native(GOLDEN_MPS / "mrpl_crude_small.mps")
    opts = smy.MipOptions(lp=smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced",
                                             halpern=True, verify=False), gap_tol=1e-4, time_limit=120.0)
    r = smy.solve_mip(md, None, opts)
    print(f"mrpl_crude_small: {r.objective!r} (HiGHS {want!r}), bound {r.bound!r}, {r.nodes} nodes, "
          f"{r.waves} waves, {r.fixed_solves} fixed-integer batches, {r.seconds:.1f} s")
    assert r.status == smy.OPTIMAL, r
    assert r.objective <= want + 1e-9 * abs(want)  # a maximization: no feasible point beats it
    assert r.bound >= want - 1e-9 * abs(want)
    assert r.bound - r.objective <= 1e-4 * (1 + abs(r.objective))


def equality_mip(k: int, rhs: float):
    """min sum x s.t. 2 x_1 + ... + 2 x_k = rhs, x integer in [0, 1]."""
    return smy.Model(1, k, np.arange(k + 1, dtype=np.int64), np.zeros(k, np.int32), np.full(k, 2.0),
                     np.ones(k), np.zeros(k), np.ones(k), np.array([rhs]), np.array([rhs]), False, 0.0)


def test_integer_infeasible_mips(gpu):
    # Both have a feasible LP (x = 1/4 each, x = 1/2 each) and no integer point (HiGHS agrees,
    # golden mip_knapsack). 2 x1 + 2 x2 = 1: the rows' implied bounds already cross (x_j <= 1/2
    # rounds to 0, then x_j >= 1/2 to 1), before any relaxation. 2 x1 + 2 x2 + 2 x3 = 3: the implied
    # bounds are the given ones, and the branches end in infeasible relaxations.
    fx = load_golden("mip_knapsack")
    hand = {c["handmade"]: c["status"] for c in fx["cases"] if "handmade" in c}
    opts = smy.MipOptions(lp=smy.PdlpOptions(tol=1e-8, verify=False))
    assert hand["2 x1 + 2 x2 = 1"] == smy.INFEASIBLE
    r = smy.solve_mip(equality_mip(2, 1.0), [True] * 2, opts)
    assert r.status == smy.INFEASIBLE and r.col_value is None and r.nodes == 0, r
    assert hand["2 x1 + 2 x2 + 2 x3 = 3"] == smy.INFEASIBLE
    r = smy.solve_mip(equality_mip(3, 3.0), [True] * 3, opts)
    assert r.status == smy.INFEASIBLE and r.col_value is None and r.pruned_infeasible >= 2, r


def test_mip_refuses_a_model_without_integer_columns(gpu):
    md = G.knapsack(7001, 3, 30, 3)
    with pytest.raises(smy.NativeError):
        smy.solve_mip(md, np.zeros(md.n, bool))


def test_minimization_form(gpu):
    # The first knapsack as a minimization of the negated profits: the same search in the other
    # sense reaches minus HiGHS's optimum, with its bound (now a lower bound) below the incumbent.
    fx = load_golden("mip_knapsack")
    case = [c for c in fx["cases"] if "knapsack" in c][0]
    md = knapsack_case(case)
    mn = smy.Model(md.m, md.n, md.start, md.index, md.value, -md.obj, md.col_lower, md.col_upper,
                   md.row_lower, md.row_upper, False, 0.0)
    opts = smy.MipOptions(lp=smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced",
                                             halpern=True, verify=False), gap_tol=1e-4, wave=64)
    r = smy.solve_mip(mn, np.ones(md.n, bool), opts)
    want = -case["objective"]
    assert r.status == smy.OPTIMAL, r
    assert abs(r.objective - want) <= 1e-6 * (1 + abs(want))
    assert r.objective - 1e-4 * (1 + abs(r.objective)) - 1e-9 <= r.bound <= r.objective + 1e-9
