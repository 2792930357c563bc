"""Batched PDLP on the GPU: many LPs stacked on one device handle, stepping together.

A batch gives every LP exactly the arithmetic it gets alone (its blocks, sums and decisions do
not depend on the other instances), so the invariants are bit-level: a batch of K equals K single
solves (status, iterations, restarts, objective to the last bit), K copies of one LP give K
identical answers, and the device sums of a batched handle equal those of K single handles."""
import numpy as np
import pytest

import samaya_amd as smy
import samaya_amd.device as D
import samaya_amd.generators as G
from conftest import load_golden
from oracle_lib import M
from test_gpu_pdhg import equilibrated, identity_lp, make_device, transpose_csc

pytestmark = pytest.mark.gpu


def mixed_models():
    """LPs of uneven size and shape: sparse uniform, power-law columns, two small blending LPs."""
    return [
        G.sparse_uniform(11, 60, 150, 4),
        G.sparse_powerlaw(12, 300, 700, 40, 5.0),
        G.sparse_uniform(13, 500, 1200, 6),
        G.blending(14, periods=2, units=3, streams=4, qualities=2),
        G.sparse_uniform(15, 1, 3, 1),
        G.blending(16, periods=3, units=2, streams=5, qualities=3),
    ]


def same(a, b):
    assert a.status == b.status
    assert a.iterations == b.iterations
    assert a.fp32_iterations == b.fp32_iterations
    assert a.rounds == b.rounds
    assert a.verified == b.verified
    assert a.objective == b.objective or (np.isnan(a.objective) and np.isnan(b.objective))
    np.testing.assert_array_equal(a.col_value, b.col_value)
    np.testing.assert_array_equal(a.row_dual, b.row_dual)


@pytest.mark.parametrize("precision,algorithm", [("fp64", "reference"), ("mixed", "enhanced")])
def test_batch_equals_single_solves(gpu, precision, algorithm):
    models = mixed_models()
    opts = smy.PdlpOptions(tol=1e-6, max_iterations=50000, precision=precision, algorithm=algorithm)
    singles = [smy.solve_lp(md, opts) for md in models]
    batch, times = smy.solve_lp_batch(models, opts)
    assert len(batch) == len(models)
    for a, b in zip(batch, singles):
        same(a, b)
    assert all(r.status == smy.OPTIMAL for r in singles)
    assert times.solve_seconds >= times.setup_seconds > 0.0 and times.kernel_ms > 0.0
    if precision == "mixed":
        assert any(r.fp32_iterations > 0 for r in batch)


def test_batch_with_polish_rounds_equals_single_solves(gpu):
    # Instances that the verifier rejects go through polish rounds of their own (a batch per
    # round): the result, round count included, is the single solve's.
    models = [G.sparse_uniform(21, 400, 1000, 8), G.sparse_uniform(22, 200, 600, 5),
              G.sparse_powerlaw(23, 500, 900, 60, 6.0)]
    opts = smy.PdlpOptions(tol=1e-4, max_iterations=50000, polish_rounds=2)
    singles = [smy.solve_lp(md, opts) for md in models]
    batch, _ = smy.solve_lp_batch(models, opts)
    for a, b in zip(batch, singles):
        same(a, b)
    assert any(r.rounds > 1 for r in singles)


def test_identical_instances_give_identical_answers(gpu):
    md = G.sparse_uniform(31, 300, 800, 6)
    batch, _ = smy.solve_lp_batch([md] * 5, smy.PdlpOptions(tol=1e-6, precision="mixed",
                                                             polish_rounds=3))
    for r in batch[1:]:
        same(r, batch[0])
    assert batch[0].status == smy.OPTIMAL and batch[0].verified


def test_infeasible_and_unbounded_instances_in_a_batch(gpu):
    # Detection runs per instance: the batch stops each infeasible / unbounded LP where its own
    # solve stops, with the same certificate verdicts, while the feasible ones continue.
    fx = load_golden("pdlp_infeasible_unbounded")
    models = [M.from_json(c["model"]).product_model() for c in fx["infeasible"][:3]]
    models += [M.from_json(c["model"]).product_model() for c in fx["unbounded"][:3]]
    models.insert(2, G.sparse_uniform(41, 100, 300, 5))
    opts = smy.PdlpOptions(tol=1e-4, max_iterations=100000)
    singles = [smy.solve_lp(md, opts) for md in models]
    batch, _ = smy.solve_lp_batch(models, opts)
    for a, b in zip(batch, singles):
        same(a, b)
    assert {r.status for r in batch} >= {smy.INFEASIBLE, smy.UNBOUNDED, smy.OPTIMAL}


def _device_lp(md, seed):
    """The equilibrated LP of a model with random scales U(0.5, 2), as the device tests use."""
    lp = identity_lp(M.from_product(md))
    rng = np.random.default_rng(seed)
    rs = rng.uniform(0.5, 2.0, lp.m)
    cs = rng.uniform(0.5, 2.0, lp.n)
    return lp, equilibrated(lp, rs, cs), rs, cs


def test_batched_device_sums_equal_single_handles(gpu):
    # Device level: 64 steps (graph), the evaluation sums of both iterates, a restart of half the
    # instances, 3 more steps, distances: every number equal to a single handle's, bit for bit,
    # with one instance frozen halfway and one switched to FP32.
    models = [G.sparse_uniform(51, 80, 200, 4), G.sparse_powerlaw(52, 300, 500, 30, 4.0),
              G.blending(53, periods=2, units=2, streams=3, qualities=2), G.sparse_uniform(54, 40, 90, 3)]
    inputs = [_device_lp(md, 60 + i) for i, md in enumerate(models)]
    K = len(models)
    singles = []
    for lp, eq, rs, cs in inputs:
        dev = make_device(eq)
        dev.set_original(cs, rs, lp.cost, lp.lower, lp.upper)
        singles.append(dev)
    batch = D.PdhgBatch.create([(eq.m, eq.n, eq.start, eq.index, eq.value) for _, eq, _, _ in inputs],
                               [transpose_csc(eq) for _, eq, _, _ in inputs],
                               [eq.cost for _, eq, _, _ in inputs], [eq.lower for _, eq, _, _ in inputs],
                               [eq.upper for _, eq, _, _ in inputs])
    assert batch.size == K
    for i, (lp, eq, rs, cs) in enumerate(inputs):
        batch.set_original(i, cs, rs, lp.cost, lp.lower, lp.upper)
    tau = [0.3, 0.2, 0.25, 0.4]
    sigma = [0.3, 0.45, 0.35, 0.2]

    for i, dev in enumerate(singles):
        dev.run(tau[i], sigma[i], 1.0, 0, 64)
    batch.run(tau, sigma, 1.0, [0] * K, 64)
    got = batch.evaluate()
    for i, dev in enumerate(singles):
        np.testing.assert_array_equal(got[i], dev.evaluate_both())

    which = [D.AVERAGE, -1, D.CURRENT, -1]
    dist = batch.restart_distance(which)
    for i, dev in enumerate(singles):
        if which[i] >= 0:
            np.testing.assert_array_equal(dist[i], dev.restart_distance(which[i]))
    batch.restart_to(which)
    for i, dev in enumerate(singles):
        if which[i] >= 0:
            dev.restart_to(which[i])
    batch.advance()
    for dev in singles:
        dev.advance()
    batch.freeze(3)
    batch.set_precision([-1, "fp32", -1, -1])
    singles[1].set_precision("fp32")
    k0 = [0, 64, 0, 0]
    for i, dev in enumerate(singles[:3]):
        dev.run(tau[i], sigma[i], 1.0, k0[i], 3)
    batch.run(tau, sigma, 1.0, k0, 3)
    got = batch.evaluate()
    for i, dev in enumerate(singles[:3]):
        np.testing.assert_array_equal(got[i], dev.evaluate_both())
        mine, theirs = batch.download(i), dev.download()
        for key in theirs:
            np.testing.assert_array_equal(mine[key], theirs[key])
