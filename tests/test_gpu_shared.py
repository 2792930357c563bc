"""The shared-matrix batch on the GPU: K LPs with one matrix and cost and their own column bounds
(branch-and-bound node relaxations), iterates stored instance-minor and products as SpMMs.

It sums in another order than the stacked batch, so the two agree to rounding, not to the bit:
device sums after the same steps agree to 1e-9 relative (FP64), answers agree in status and
objective to the solve tolerance. Identical instances give identical bits, and node relaxations
match solve_lp of the model with those bounds."""
import numpy as np
import pytest

from conftest import load_golden

import samaya_amd as smy
import samaya_amd.device as D
import samaya_amd.generators as G
from oracle_lib import M
from test_gpu_pdhg import equilibrated, identity_lp, transpose_csc

pytestmark = pytest.mark.gpu


def node_bounds(md, count, seed0=900, lo_frac=0.01, hi_frac=0.05):
    out = [G.fixings(seed0 + k, md.col_lower, md.col_upper, lo_frac, hi_frac) for k in range(count)]
    return [o[0] for o in out], [o[1] for o in out]


def rel_close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= tol * (1.0 + np.abs(b)))


@pytest.mark.parametrize("K", [7, 72])  # few instances: per-(row, instance) steps; many: SpMM rows
def test_shared_device_matches_stacked_device(gpu, K):
    md = G.knapsack(71, 60, 150, 4)
    lp = identity_lp(M.from_product(md))
    rng = np.random.default_rng(3)
    rs, cs = rng.uniform(0.5, 2.0, lp.m), rng.uniform(0.5, 2.0, lp.n)
    eq = equilibrated(lp, rs, cs)
    lows, ups = [], []
    for k in range(K):
        lo, up = G.fixings(800 + k, eq.lower[:lp.n], eq.upper[:lp.n], 0.05, 0.2)
        lows.append(np.concatenate([lo, eq.lower[lp.n:]]))
        ups.append(np.concatenate([up, eq.upper[lp.n:]]))
    a = (eq.m, eq.n, eq.start, eq.index, eq.value)
    at = transpose_csc(eq)
    shared = D.PdhgBatch.create_shared(a, at, eq.cost, lows, ups)
    stacked = D.PdhgBatch.create([a] * K, [at] * K, [eq.cost] * K, lows, ups)
    ol = [np.concatenate([lows[k][:lp.n] * cs, lp.lower[lp.n:]]) for k in range(K)]
    ou = [np.concatenate([ups[k][:lp.n] * cs, lp.upper[lp.n:]]) for k in range(K)]
    for b in (shared, stacked):
        for k in range(K):
            b.set_original(k, cs, rs, lp.cost, ol[k], ou[k])
    tau = 0.2 + 0.2 * (np.arange(K) % 7) / 6
    sigma = 0.4 - 0.2 * (np.arange(K) % 5) / 4
    for b in (shared, stacked):
        b.run(tau, sigma, 1.0, [0] * K, 64)
    assert rel_close(shared.evaluate(), stacked.evaluate(), 1e-9)
    which = [(0, -1, 1, -1, 0, 1, -1)[k % 7] for k in range(K)]
    flagged = [k for k in range(K) if which[k] >= 0]  # the others' distances are not computed
    assert rel_close(shared.restart_distance(which)[flagged], stacked.restart_distance(which)[flagged], 1e-9)
    k0 = [(0 if which[k] >= 0 else 64) for k in range(K)]
    for b in (shared, stacked):
        b.restart_to(which)
        b.advance()
        b.freeze(3)
        b.set_precision(["fp32" if k == 1 else -1 for k in range(K)])
        b.run(tau, sigma, 1.0, k0, 5)
    es, et = shared.evaluate(), stacked.evaluate()
    for k in range(K):
        if k not in (1, 3):
            assert rel_close(es[k], et[k], 1e-9), k
    assert rel_close(es[1], et[1], 1e-4)  # FP32 steps: rounding of FP32 sums in another order
    for k in (0, 2, 5):
        ds, dt = shared.download(k), stacked.download(k)
        for key in ds:
            assert rel_close(ds[key], dt[key], 1e-9), key


def test_coded_bounds_give_the_same_bits(gpu):
    # Node bounds (the root's or a fixing) have at most 4 (lower, upper) pairs per column, so the
    # shared batch keeps them as one-byte codes. Four more instances with pairs of their own in
    # column 0 make it keep the full bound arrays instead. The 72 common instances step to the same
    # bits either way, in FP64 and in FP32.
    md = G.knapsack(74, 60, 150, 4)
    lp = identity_lp(M.from_product(md))
    rng = np.random.default_rng(6)
    rs, cs = rng.uniform(0.5, 2.0, lp.m), rng.uniform(0.5, 2.0, lp.n)
    eq = equilibrated(lp, rs, cs)
    K = 72
    lows, ups = [], []
    for k in range(K):
        lo, up = G.fixings(820 + k, eq.lower[:lp.n], eq.upper[:lp.n], 0.05, 0.2)
        lows.append(np.concatenate([lo, eq.lower[lp.n:]]))
        ups.append(np.concatenate([up, eq.upper[lp.n:]]))
    extra_lo, extra_up = [], []
    for e in range(4):
        lo, up = lows[0].copy(), ups[0].copy()
        lo[0], up[0] = 0.01 * (e + 1), 0.5 + 0.01 * e
        extra_lo.append(lo)
        extra_up.append(up)
    a = (eq.m, eq.n, eq.start, eq.index, eq.value)
    at = transpose_csc(eq)
    coded = D.PdhgBatch.create_shared(a, at, eq.cost, lows, ups)
    full = D.PdhgBatch.create_shared(a, at, eq.cost, lows + extra_lo, ups + extra_up)
    ol = [np.concatenate([x[:lp.n] * cs, lp.lower[lp.n:]]) for x in lows + extra_lo]
    ou = [np.concatenate([x[:lp.n] * cs, lp.upper[lp.n:]]) for x in ups + extra_up]
    for b, count in ((coded, K), (full, K + 4)):
        for k in range(count):
            b.set_original(k, cs, rs, lp.cost, ol[k], ou[k])
    tau = 0.2 + 0.2 * (np.arange(K + 4) % 7) / 6
    sigma = 0.4 - 0.2 * (np.arange(K + 4) % 5) / 4
    coded.run(tau[:K], sigma[:K], 1.0, [0] * K, 64)
    full.run(tau, sigma, 1.0, [0] * (K + 4), 64)
    np.testing.assert_array_equal(coded.evaluate(), full.evaluate()[:K])
    for b, count in ((coded, K), (full, K + 4)):
        b.advance()
        b.set_precision(["fp32"] * count)
    coded.run(tau[:K], sigma[:K], 1.0, [64] * K, 64)
    full.run(tau, sigma, 1.0, [64] * (K + 4), 64)
    np.testing.assert_array_equal(coded.evaluate(), full.evaluate()[:K])
    for k in (0, 17, 71):
        dc, df = coded.download(k), full.download(k)
        for key in dc:
            np.testing.assert_array_equal(dc[key], df[key], err_msg=key)


def test_identical_bound_vectors_give_identical_answers(gpu):
    md = G.knapsack(72, 300, 800, 6)
    lo, up = G.fixings(901, md.col_lower, md.col_upper)
    res, _ = smy.solve_lp_bounds_batch(md, [lo] * 6, [up] * 6,
                                       smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced"))
    for r in res[1:]:
        assert r.status == res[0].status and r.iterations == res[0].iterations
        assert r.objective == res[0].objective
        np.testing.assert_array_equal(r.col_value, res[0].col_value)
    assert res[0].status == smy.OPTIMAL


@pytest.mark.parametrize("precision,algorithm", [("fp64", "reference"), ("mixed", "enhanced")])
def test_node_relaxations_match_single_solves(gpu, precision, algorithm):
    md = G.knapsack(73, 400, 1000, 6)
    lows, ups = node_bounds(md, 9)
    highs = {(e["knapsack"][0], e["fixings_seed"]): e["objective"]
             for e in load_golden("node_relaxations")["nodes"] if e["status"] == 0}
    opts = smy.PdlpOptions(tol=1e-7, precision=precision, algorithm=algorithm, polish_rounds=2)
    batch, times = smy.solve_lp_bounds_batch(md, lows, ups, opts)
    assert times.kernel_ms > 0
    for k in range(len(lows)):
        mk = smy.Model(md.m, md.n, md.start, md.index, md.value, md.obj, lows[k], ups[k],
                       md.row_lower, md.row_upper, md.maximize, md.offset)
        single = smy.solve_lp(mk, opts)
        assert batch[k].status == single.status == smy.OPTIMAL
        assert abs(batch[k].objective - single.objective) <= 1e-6 * (1 + abs(single.objective))
        assert batch[k].verified == single.verified
        # HiGHS's objective of the same node (golden node_relaxations): within 1e-6 once polished.
        want = highs[(73, 900 + k)]
        assert batch[k].verified and abs(batch[k].objective - want) <= 1e-6 * (1 + abs(want))
        # The node's bound for branch and bound (a maximization: an upper bound) from its duals
        # never falls below HiGHS's optimum and is tight to 1e-6.
        assert want <= batch[k].dual_bound <= want + 1e-6 * (1 + abs(want))
        # The answer respects the node's fixings.
        fixed = lows[k] == ups[k]
        assert np.max(np.abs(batch[k].col_value[fixed] - lows[k][fixed])) <= 1e-6


def test_shared_and_stacked_batches_agree(gpu):
    md = G.knapsack(74, 300, 700, 5)
    lows, ups = node_bounds(md, 5)
    opts = smy.PdlpOptions(tol=1e-7, precision="mixed", algorithm="enhanced")
    shared, _ = smy.solve_lp_bounds_batch(md, lows, ups, opts)
    opts.share_matrix = False
    stacked, _ = smy.solve_lp_bounds_batch(md, lows, ups, opts)
    for a, b in zip(shared, stacked):
        assert a.status == b.status == smy.OPTIMAL
        assert abs(a.objective - b.objective) <= 1e-6 * (1 + abs(b.objective))


def test_warm_start_is_used(gpu):
    # Nodes started from their own tighter answers, with their primal weight, stop at the first
    # evaluations; started from the root's answer they reach the cold starts' objectives.
    md = G.knapsack(75, 400, 1000, 6)
    opts = smy.PdlpOptions(tol=1e-6, precision="fp64", verify=False)
    lows, ups = node_bounds(md, 6)
    tight, _ = smy.solve_lp_bounds_batch(md, lows, ups, smy.PdlpOptions(tol=1e-9, verify=False))
    for k, c in enumerate(tight):
        warm = smy.PdlpOptions(tol=1e-6, precision="fp64", verify=False, start_omega=c.omega)
        again, _ = smy.solve_lp_bounds_batch(md, [lows[k]], [ups[k]], warm,
                                             start=(c.col_value, c.row_dual))
        assert again[0].status == smy.OPTIMAL and again[0].iterations <= 2 * 64, again[0].iterations
        assert abs(again[0].objective - c.objective) <= 1e-6 * (1 + abs(c.objective))
    cold, _ = smy.solve_lp_bounds_batch(md, lows, ups, opts)
    root = smy.solve_lp(md, opts)
    opts.start_omega = root.omega
    warm, _ = smy.solve_lp_bounds_batch(md, lows, ups, opts, start=(root.col_value, root.row_dual))
    for c, w in zip(cold, warm):
        assert c.status == w.status == smy.OPTIMAL
        assert abs(c.objective - w.objective) <= 1e-4 * (1 + abs(c.objective))


@pytest.mark.parametrize("K,rho", [(7, 1.0), (72, 0.5)])  # compact block / full-width SpMM rows
def test_shared_halpern_steps_match_stacked(gpu, K, rho):
    # Restarted Halpern steps (T(z) in the averages' slot, z' in the iterates) on both layouts:
    # evaluation sums, fixed-point residuals, restart distances and the iterates agree to 1e-9 in
    # FP64, through a restart at T(z), a frozen instance and an FP32 instance.
    md = G.knapsack(71, 60, 150, 4)
    lp = identity_lp(M.from_product(md))
    rng = np.random.default_rng(5)
    rs, cs = rng.uniform(0.5, 2.0, lp.m), rng.uniform(0.5, 2.0, lp.n)
    eq = equilibrated(lp, rs, cs)
    lows, ups = [], []
    for k in range(K):
        lo, up = G.fixings(850 + k, eq.lower[:lp.n], eq.upper[:lp.n], 0.05, 0.2)
        lows.append(np.concatenate([lo, eq.lower[lp.n:]]))
        ups.append(np.concatenate([up, eq.upper[lp.n:]]))
    a = (eq.m, eq.n, eq.start, eq.index, eq.value)
    at = transpose_csc(eq)
    shared = D.PdhgBatch.create_shared(a, at, eq.cost, lows, ups)
    stacked = D.PdhgBatch.create([a] * K, [at] * K, [eq.cost] * K, lows, ups)
    ol = [np.concatenate([lows[k][:lp.n] * cs, lp.lower[lp.n:]]) for k in range(K)]
    ou = [np.concatenate([ups[k][:lp.n] * cs, lp.upper[lp.n:]]) for k in range(K)]
    for b in (shared, stacked):
        b.set_halpern(True, rho)
        for k in range(K):
            b.set_original(k, cs, rs, lp.cost, ol[k], ou[k])
    tau = 0.2 + 0.2 * (np.arange(K) % 7) / 6
    sigma = 0.4 - 0.2 * (np.arange(K) % 5) / 4
    for b in (shared, stacked):
        b.run(tau, sigma, 1.0, [0] * K, 64)
    es, et = shared.evaluate(), stacked.evaluate()
    assert rel_close(es, et, 1e-9)
    np.testing.assert_array_equal(es[:, :7], es[:, 7:])  # T(z) in both halves
    assert rel_close(shared.fixed_point(), stacked.fixed_point(), 1e-9)
    which = [(1, -1, 1, -1, 1, 1, -1)[k % 7] for k in range(K)]
    flagged = [k for k in range(K) if which[k] >= 0]
    assert rel_close(shared.restart_distance(which)[flagged], stacked.restart_distance(which)[flagged], 1e-9)
    k0 = [(0 if which[k] >= 0 else 64) for k in range(K)]
    for b in (shared, stacked):
        b.restart_to(which)
        b.advance()
        b.freeze(3)
        b.set_precision(["fp32" if k == 1 else -1 for k in range(K)])
        b.run(tau, sigma, 1.0, k0, 5)
    es, et = shared.evaluate(), stacked.evaluate()
    fs, ft = shared.fixed_point(), stacked.fixed_point()
    for k in range(K):
        if k not in (1, 3):
            assert rel_close(es[k], et[k], 1e-9), k
            assert rel_close(fs[k], ft[k], 1e-9), k
    assert rel_close(es[1], et[1], 1e-4)  # FP32 steps: rounding of FP32 sums in another order
    for k in (0, 2, 5):
        ds, dt = shared.download(k), stacked.download(k)
        for key in ds:
            assert rel_close(ds[key], dt[key], 1e-9), key


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_shared_halpern_node_relaxations(gpu, precision):
    # Halpern steps on the shared-matrix batch: every node as the stacked batch solves it, and
    # HiGHS's objective of the node (golden node_relaxations) once polished.
    md = G.knapsack(73, 400, 1000, 6)
    lows, ups = node_bounds(md, 9)
    highs = {(e["knapsack"][0], e["fixings_seed"]): e["objective"]
             for e in load_golden("node_relaxations")["nodes"] if e["status"] == 0}
    opts = smy.PdlpOptions(tol=1e-7, precision=precision, algorithm="enhanced", halpern=True,
                           polish_rounds=2)
    shared, times = smy.solve_lp_bounds_batch(md, lows, ups, opts)
    assert times.kernel_ms > 0
    opts.share_matrix = False
    stacked, _ = smy.solve_lp_bounds_batch(md, lows, ups, opts)
    for k in range(len(lows)):
        a, b = shared[k], stacked[k]
        assert a.status == b.status == smy.OPTIMAL
        assert abs(a.objective - b.objective) <= 1e-6 * (1 + abs(b.objective))
        want = highs[(73, 900 + k)]
        assert a.verified and abs(a.objective - want) <= 1e-6 * (1 + abs(want))
        assert want <= a.dual_bound <= want + 1e-6 * (1 + abs(want))
