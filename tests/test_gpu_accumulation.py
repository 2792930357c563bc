"""FP32 steps with FP64 accumulation (smy_pdhg_set_fp64_accumulation; the solver's
fp64_accumulation). The FP32 step kernels sum each product in FP64 and form the update that uses
it in FP64, rounding once to FP32:
- their updates lie closer to the FP64 step's than FP32 accumulation's do;
- a batch still equals its single handles bit for bit;
- solves reach the same answers as FP64 PDLP."""
import numpy as np
import pytest

import samaya_amd as smy
import samaya_amd.device as D
import samaya_amd.generators as G
from test_gpu_batch import _device_lp
from test_gpu_pdhg import transpose_csc

pytestmark = pytest.mark.gpu


def make_batch(inputs, precision, acc64):
    b = D.PdhgBatch.create([(eq.m, eq.n, eq.start, eq.index, eq.value) for _, eq, _, _ in inputs],
                           [transpose_csc(eq) for _, eq, _, _ in inputs],
                           [eq.cost for _, eq, _, _ in inputs], [eq.lower for _, eq, _, _ in inputs],
                           [eq.upper for _, eq, _, _ in inputs], precision)
    for i, (lp, eq, rs, cs) in enumerate(inputs):
        b.set_original(i, cs, rs, lp.cost, lp.lower, lp.upper)
    if acc64:
        b.set_fp64_accumulation(True)
    return b


def models():
    # Long rows and columns (up to 2,000 nonzeros), where FP32 sums lose the most.
    return [G.sparse_powerlaw(201, 2000, 4000, 2000, 30.0), G.sparse_uniform(202, 300, 700, 25)]


def test_updates_are_closer_to_fp64(gpu):
    inputs = [_device_lp(md, 210 + i) for i, md in enumerate(models())]
    K = len(inputs)
    rng = np.random.default_rng(7)
    starts = [(rng.uniform(0.0, 1.0, eq.n + eq.m), rng.normal(0.0, 1.0, eq.m)) for _, eq, _, _ in inputs]
    out = {}
    for name, precision, acc64 in (("fp64", "fp64", False), ("acc32", "fp32", False), ("acc64", "fp32", True)):
        b = make_batch(inputs, precision, acc64)
        for i, (v, y) in enumerate(starts):
            b.set_start(i, v, y)
        b.run([0.01] * K, [0.01] * K, 1.0, [0] * K, 1)
        out[name] = [b.download(i) for i in range(K)]
    for i in range(K):
        ref = out["fp64"][i]
        for key in ("y_new", "v_new"):
            e32 = np.max(np.abs(out["acc32"][i][key] - ref[key]))
            e64 = np.max(np.abs(out["acc64"][i][key] - ref[key]))
            assert e64 <= 1.01 * e32, (i, key, e64, e32)
        assert np.any(out["acc32"][i]["y_new"] != out["acc64"][i]["y_new"])  # the option acts


@pytest.mark.parametrize("adaptive", [False, True])
def test_batch_equals_single_handles(gpu, adaptive):
    inputs = [_device_lp(md, 220 + i) for i, md in enumerate(models())]
    K = len(inputs)

    def steps(b, k):
        if adaptive:
            st = D.PdhgBatch.adaptive_state([0.05] * k, 1.0)
            b.run_adaptive(st, 64)
        else:
            b.run([0.05] * k, [0.05] * k, 1.0, [0] * k, 64)
        return b.evaluate()

    batch = steps(make_batch(inputs, "fp32", True), K)
    for i in range(K):
        single = steps(make_batch([inputs[i]], "fp32", True), 1)
        np.testing.assert_array_equal(batch[i], single[0])


@pytest.mark.parametrize("adaptive", [False, True])
def test_solves_agree_with_fp64(gpu, adaptive):
    md = G.sparse_powerlaw(161, 1500, 4000, 300, 9.0)
    lp = smy.make_problem(md)
    ref = smy.solve_pdlp(lp, smy.PdlpOptions(tol=1e-6, precision="fp64"))
    got = smy.solve_pdlp(lp, smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced",
                                              fp64_accumulation=True, adaptive_steps=adaptive))
    assert ref.status == got.status == smy.OPTIMAL
    # Both answers meet rel KKT 1e-6, at different points: the primal residual the tolerance
    # allows moves the objective by ~2e-5 relative here.
    obj_ref, obj = float(lp.cost @ ref.x), float(lp.cost @ got.x)
    assert abs(obj - obj_ref) <= 1e-4 * (1 + abs(obj_ref))
