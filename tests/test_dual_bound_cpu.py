"""The weak-duality bound from row duals (smy_lp_dual_bound), pinned by what duality fixes: at the
optimal duals of a tiny LP it equals the optimum (tests/instances/tiny_lp.mps: max 3x + 2y,
x + y <= 4, x + 3y <= 6, 0 <= x <= 3, y >= 0, optimum 11 at (3, 1) with duals (2, 0)); at any other
duals it never crosses the optimum (HiGHS's, on random LPs of the reference's families); a free
column with a nonzero reduced cost gives no bound."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O
import samaya_amd as smy
from test_oracle import highs_objective, tiny_lp_model


def product_model(md: O.M):
    return smy.Model(md.m, md.n, md.start.astype(np.int64), md.index.astype(np.int32),
                     md.value.astype(float), md.obj, md.col_lower, md.col_upper, md.row_lower,
                     md.row_upper, md.maximize, md.offset)


def test_tiny_lp_bound_is_tight_at_its_duals():
    md = product_model(tiny_lp_model())
    # Row duals in the d = c - A'y convention: d = (3 - 2, 2 - 2) = (1, 0), x at its upper 3.
    assert smy.dual_bound(md, np.array([2.0, 0.0])) == pytest.approx(11.0, abs=1e-12)
    # Any other sign-feasible duals give an upper bound above 11 (maximization).
    rng = np.random.default_rng(3)
    for _ in range(50):
        y = rng.uniform(0, 5, 2)
        assert smy.dual_bound(md, y) >= 11.0 - 1e-12
    # Sign-infeasible duals are projected first (a <= row's dual cannot be negative here).
    assert smy.dual_bound(md, np.array([2.0, -7.0])) == pytest.approx(11.0, abs=1e-12)


def test_bound_never_crosses_the_optimum():
    checked = 0
    for family in (0, 4):  # feasible and boxed families (tests/lp_generators.hpp)
        rng = O.RefRng(77 + family) if O.ref_available() else None
        if rng is None:
            pytest.skip("needs the reference build for its generator")
        for _ in range(12):
            md = rng.random_lp(family, 12, 12)
            try:
                opt = highs_objective(md)
            except AssertionError:
                continue
            pm = product_model(md)
            r = np.random.default_rng(checked)
            for _ in range(5):
                y = r.normal(size=md.m) * 3
                b = smy.dual_bound(pm, y)
                if md.maximize:
                    assert b >= opt - 1e-9 * (1 + abs(opt))
                else:
                    assert b <= opt + 1e-9 * (1 + abs(opt))
            checked += 1
    assert checked >= 10


def test_free_column_with_reduced_cost_gives_no_bound():
    A = sp.csc_matrix(np.array([[1.0, 1.0]]))
    md = smy.Model(1, 2, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data,
                   np.array([1.0, 2.0]), np.array([-np.inf, 0.0]), np.array([np.inf, 5.0]),
                   np.array([1.0]), np.array([1.0]))
    assert smy.dual_bound(md, np.array([0.5])) == -np.inf   # d_0 = 0.5 on a free column
    assert smy.dual_bound(md, np.array([1.0])) == pytest.approx(1.0, abs=1e-12)  # d = (0, 1)
