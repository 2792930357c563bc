"""Model-level solves with the columns reordered (PdlpOptions.column_order: by decreasing nonzero
count, or the touch order: hot columns by count, the rest by the first row holding them): the
answer comes back in the model's order and is exactly the answer of the explicitly permuted model
(for the touch order, permuted by a NumPy restatement of the rule); node batches permute their
bounds and warm start alike; "auto" takes the touch order for skewed column counts (power-law LPs)
and keeps the model's otherwise."""
import numpy as np
import pytest
import scipy.sparse as sp

import samaya_amd as smy

pytestmark = pytest.mark.gpu
G = smy.generators


def permuted(md, order):
    A = sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))[:, order]
    A.sort_indices()
    return smy.Model(md.m, md.n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data,
                     md.obj[order], md.col_lower[order], md.col_upper[order], md.row_lower,
                     md.row_upper, md.maximize, md.offset)


def degree_order(md):
    return np.argsort(-np.diff(md.start), kind="stable")


def touch_order(md):
    """Hot columns (count >= 2x the mean) by decreasing count, then the others by their first row
    (ties in the model's order)."""
    counts = np.diff(md.start)
    first = np.array([md.index[md.start[j]:md.start[j + 1]].min() if counts[j] else md.m
                      for j in range(md.n)])
    hot = counts >= 2.0 * md.nnz / md.n
    h, c = np.flatnonzero(hot), np.flatnonzero(~hot)
    return np.concatenate([h[np.argsort(-counts[h], kind="stable")], c[np.argsort(first[c], kind="stable")]])


def test_degree_order_equals_the_permuted_model(gpu):
    md = G.c2(0.01)
    order = degree_order(md)
    opts = dict(tol=1e-6, precision="mixed", algorithm="enhanced", adaptive_steps=True)
    got = smy.solve_lp(md, smy.PdlpOptions(**opts, column_order="degree"))
    ref = smy.solve_lp(permuted(md, order), smy.PdlpOptions(**opts, column_order="natural"))
    assert got.status == ref.status == smy.OPTIMAL
    assert got.iterations == ref.iterations and got.objective == ref.objective
    x = np.empty(md.n)
    x[order] = ref.col_value
    d = np.empty(md.n)
    d[order] = ref.col_dual
    assert np.array_equal(got.col_value, x) and np.array_equal(got.col_dual, d)
    assert np.array_equal(got.row_dual, ref.row_dual)
    # The answer in the model's order is the model's: its verifier numbers are the permuted one's.
    v = smy.verify_lp_optimality(md, got.col_value, got.row_dual)
    assert abs(v[1] - got.max_primal_violation) <= 1e-12 * (1 + got.max_primal_violation)
    # "auto" takes the touch order on this power-law LP and the model's order on a uniform one.
    auto = smy.solve_lp(md, smy.PdlpOptions(**opts))
    touch = smy.solve_lp(md, smy.PdlpOptions(**opts, column_order="touch"))
    assert auto.iterations == touch.iterations and auto.objective == touch.objective
    uni = G.sparse_uniform(5, 300, 800, 5)
    a = smy.solve_lp(uni, smy.PdlpOptions(**opts))
    b = smy.solve_lp(uni, smy.PdlpOptions(**opts, column_order="natural"))
    assert a.iterations == b.iterations and a.objective == b.objective


def test_node_batch_in_degree_order_equals_single_solves(gpu):
    md = G.sparse_powerlaw(31, 300, 900, 120, 4.0)
    rng = np.random.default_rng(4)
    lows, ups = [], []
    for k in range(4):
        lo, up = md.col_lower.copy(), md.col_upper.copy()
        fix = rng.choice(md.n, 20, replace=False)
        up[fix] = 0.0
        lows.append(lo)
        ups.append(up)
    opts = smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced", column_order="degree",
                           adaptive_steps=True)
    batch, _ = smy.solve_lp_bounds_batch(md, lows, ups, opts)
    for k in range(4):
        node = smy.Model(md.m, md.n, md.start, md.index, md.value, md.obj, lows[k], ups[k],
                         md.row_lower, md.row_upper, md.maximize, md.offset)
        one = smy.solve_lp(node, opts)
        assert batch[k].status == one.status
        assert batch[k].iterations == one.iterations and batch[k].objective == one.objective
        assert np.array_equal(batch[k].col_value, one.col_value)
        fixed = ups[k] == 0.0
        assert np.all(batch[k].col_value[fixed] == 0.0)


def test_touch_order_equals_the_permuted_model(gpu):
    md = G.c2(0.01)
    order = touch_order(md)
    assert not np.array_equal(order, degree_order(md))
    opts = dict(tol=1e-6, precision="mixed", algorithm="enhanced", halpern=True)
    got = smy.solve_lp(md, smy.PdlpOptions(**opts, column_order="touch"))
    ref = smy.solve_lp(permuted(md, order), smy.PdlpOptions(**opts, column_order="natural"))
    assert got.status == ref.status == smy.OPTIMAL
    assert got.iterations == ref.iterations and got.objective == ref.objective
    x = np.empty(md.n)
    x[order] = ref.col_value
    assert np.array_equal(got.col_value, x) and np.array_equal(got.row_dual, ref.row_dual)
