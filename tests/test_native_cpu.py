"""CPU-side checks of the native library: it loads and exports every symbol the C headers
declare; host-side pieces (generators, scaling, computational form, verifier) agree with the
oracle; GPU entry points fail loudly without a GPU (no CPU fallback)."""
import hashlib
import math
import re
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
import samaya_amd as smy
from conftest import load_golden
from samaya_amd import _native as N

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    syms = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(smy_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    lib = smy.load()
    declared = header_symbols()
    assert len(declared) >= 40
    for s in sorted(declared):
        assert hasattr(lib, s), f"{s} declared in include/ but not exported"
    assert set(N.EXPORTS) == declared


def test_threads_per_row_known_answers():
    # tests/test_gpu.cpp:750-751.
    from samaya_amd.device import spmv_threads_per_row
    assert spmv_threads_per_row(4117330, 467376) == 16
    assert spmv_threads_per_row(4117330, 1001520) == 8
    assert spmv_threads_per_row(0, 0) == 2
    assert spmv_threads_per_row(10**9, 10) == 32


def test_generators_are_deterministic_and_pinned():
    a, b = smy.generators.c1(), smy.generators.c1()
    assert np.array_equal(a.value, b.value) and np.array_equal(a.index, b.index)
    assert hashlib.sha256(a.value.tobytes()).hexdigest() == load_golden("c1_reference")["model_sha"]
    assert a.m == 2000 and a.n == 5000 and a.nnz == 50000
    counts = np.diff(a.start)
    assert np.all(counts == 10)
    # Distinct sorted rows within every column; feasible by construction around x0.
    for j in range(a.n):
        rows = a.index[a.start[j]:a.start[j + 1]]
        assert np.all(np.diff(rows) > 0)
    assert np.all(a.row_lower[0::2] == a.row_upper[0::2])
    assert np.all(np.isneginf(a.row_lower[1::2]))


def test_powerlaw_generator_shape():
    md = smy.generators.c2(scale=0.02)  # 10,000 x 20,000, same structure
    counts = np.diff(md.start)
    assert counts.min() >= 2 and counts.max() == min(5000, md.m) and counts[0] == counts.max()
    assert abs(counts.mean() - 12.0) < 1.0
    mags = np.abs(md.value)
    assert mags.max() / mags.min() > 1e6  # values spread over 10^[-2, 2] x N(0, 1)


@pytest.mark.parametrize("name", ["pdlp_random_feasible", "pdlp_infeasible_unbounded"])
def test_scaling_and_computational_form_match_oracle(name):
    fx = load_golden(name)
    cases = fx["cases"] if "cases" in fx else fx["infeasible"] + fx["unbounded"]
    for c in cases:
        md = O.M.from_json(c["model"])
        lp, col, row = O.make_problem(md)
        pm = md.product_model()
        cs, rs = smy.compute_scaling(pm)
        assert np.array_equal(cs, col) and np.array_equal(rs, row)
        plp = smy.make_problem(pm)
        for k in ("value", "cost", "lower", "upper"):
            assert np.array_equal(getattr(plp, k), getattr(lp, k)), k


def test_c1_scaling_matches_oracle():
    md = smy.generators.c1()
    lp, col, row = O.make_problem(O.M.from_product(md))
    cs, rs = smy.compute_scaling(md)
    assert np.array_equal(cs, col) and np.array_equal(rs, row)


def test_verifier_agrees_with_oracle_on_reference_answers():
    for name in ("pdlp_random_feasible", "pdlp_through_solver"):
        for c in load_golden(name)["cases"]:
            if c["status"] != 0:
                continue
            md = O.M.from_json(c["model"])
            x = np.array([float(t) for t in c["col_value"]])
            y = np.array([float(t) for t in c["row_dual"]])
            ok_o, viol_o, _ = O.oracle_verify_lp(md, x, y)
            ok_p, pv, dv, obj, _ = smy.verify_lp_optimality(md.product_model(), x, y)
            assert ok_o == ok_p == c["verified"]
            assert pv == viol_o[0] and dv == viol_o[1] and obj == viol_o[2]


def test_verifier_certificates():
    fx = load_golden("pdlp_infeasible_unbounded")
    for c in fx["infeasible"]:
        if "ray" in c:
            ok, msg = smy.verify_infeasibility(O.M.from_json(c["model"]).product_model(),
                                               [float(t) for t in c["ray"]])
            assert ok == c["verified"], msg
    for c in fx["unbounded"]:
        if "ray" in c:
            ok, msg = smy.verify_unbounded_ray(O.M.from_json(c["model"]).product_model(),
                                               [float(t) for t in c["ray"]])
            assert ok == c["verified"], msg
    # A planted mistake: no certificate exists for a feasible LP, so any vector is rejected.
    c = next(c for c in load_golden("pdlp_random_feasible")["cases"] if c["status"] == 0)
    md = O.M.from_json(c["model"]).product_model()
    for y in ([float(t) for t in c["row_dual"]], np.ones(md.m), -np.ones(md.m)):
        ok, _ = smy.verify_infeasibility(md, y)
        assert not ok


def test_model_round_trip_and_validation():
    md = smy.generators.sparse_uniform(3, 20, 30, 4)
    nm = md.native()
    back = smy.Model._export(nm.h)
    for k in ("start", "index", "value", "obj", "col_lower", "col_upper", "row_lower", "row_upper"):
        assert np.array_equal(getattr(back, k), getattr(md, k))
    bad = smy.Model(md.m, md.n, md.start, md.index[::-1].copy(), md.value, md.obj, md.col_lower,
                    md.col_upper, md.row_lower, md.row_upper)
    with pytest.raises(smy.NativeError):
        bad.native()


@pytest.mark.skipif(smy.gpu_available(), reason="checks the behaviour without a GPU")
def test_gpu_paths_fail_loudly_without_gpu():
    lp, _, _ = O.make_problem(O.M.from_product(smy.generators.sparse_uniform(5, 10, 20, 3)))
    r = smy.solve_pdlp(lp.product(), smy.PdlpOptions(tol=1e-4))
    assert r.status == smy.NUMERICAL_ERROR and "no usable GPU" in r.message
    res = smy.solve_lp(smy.generators.sparse_uniform(5, 10, 20, 3))
    assert res.status == smy.NUMERICAL_ERROR and not res.verified
    st, obj = smy.solve(smy.generators.sparse_uniform(5, 10, 20, 3))
    assert st == 8 and math.isnan(obj)
    from samaya_amd.device import PdhgDevice
    assert PdhgDevice.create((1, 1, np.array([0, 1]), np.array([0], np.int32), np.array([1.0])),
                             (1, 1, np.array([0, 1]), np.array([0], np.int32), np.array([1.0])),
                             np.zeros(2), np.zeros(2), np.ones(2)) is None
    built, dev = smy.gpu_info()
    assert built and dev == ""
    batch, _ = smy.solve_lp_batch([smy.generators.sparse_uniform(5, 10, 20, 3)] * 2)
    assert all(r.status == smy.NUMERICAL_ERROR and "no usable GPU" in r.message for r in batch)
    from samaya_amd.device import PdhgBatch
    one = (1, 1, np.array([0, 1]), np.array([0], np.int32), np.array([1.0]))
    assert PdhgBatch.create([one, one], [one, one], [np.zeros(2)] * 2, [np.zeros(2)] * 2,
                            [np.ones(2)] * 2) is None
    assert PdhgBatch.create_shared(one, one, np.zeros(2), [np.zeros(2)] * 2, [np.ones(2)] * 2) is None
    kp = smy.generators.knapsack(7, 20, 50, 3)
    lo, up = smy.generators.fixings(8, kp.col_lower, kp.col_upper)
    nodes, _ = smy.solve_lp_bounds_batch(kp, [lo, lo], [up, up])
    assert all(r.status == smy.NUMERICAL_ERROR and "no usable GPU" in r.message for r in nodes)


def test_blending_generator_structure():
    # The batched workload's LPs: 20 units x 30 stream tanks per period, periods U{4..48},
    # qualities U{1..30}; rows = periods x (600 tank balances + 20 unit balances + 20 capacities
    # + 20 x qualities), columns = periods x (3 x 600 + 20).
    sizes = set()
    for i in range(12):
        md = smy.generators.blending(4000 + i)
        per = md.n // 1820
        assert md.n == 1820 * per and 4 <= per <= 48
        q = (md.m // per - 640) // 20
        assert md.m == per * (640 + 20 * q) and 1 <= q <= 30
        sizes.add((per, q))
        # Tank balances and unit balances are equalities, the rest <=; 0 <= x <= 10.
        eq = md.row_lower == md.row_upper
        assert eq.sum() == per * 620 and np.all(np.isneginf(md.row_lower[~eq]))
        assert np.all(md.col_lower == 0) and np.all(md.col_upper == 10)
        # Quality and capacity coefficients in [0.5, 1.5]; balance coefficients +-1; the quality
        # limits on the outputs in [-1.3, -0.9].
        v = np.abs(md.value)
        assert np.all(((v >= 0.5) & (v <= 1.5)))
    assert len(sizes) == 12  # the sizes vary with the seed
    a, b = smy.generators.blending(4007), smy.generators.blending(4007)
    assert np.array_equal(a.value, b.value) and np.array_equal(a.row_upper, b.row_upper)
    small = smy.generators.blending(4000, periods=4, qualities=1)
    assert small.m == 4 * 660 and small.n == 4 * 1820
    # Inventories link consecutive periods: the period-1 tank balances touch period-0 columns.
    t1_row = 660  # first tank balance of period 1
    cols = np.repeat(np.arange(small.n), np.diff(small.start))
    touched = cols[small.index == t1_row]
    assert touched.min() < 1820 <= touched.max()


def test_knapsack_and_fixings_generators():
    md = smy.generators.knapsack(5000, 200, 500, 8)
    assert md.maximize and md.m == 200 and md.n == 500 and md.nnz == 4000
    assert np.all(np.diff(md.start) == 8)
    assert np.all((md.value >= 1) & (md.value <= 100) & (md.value == np.round(md.value)))
    assert np.all((md.obj >= 1) & (md.obj <= 100))
    rows = np.zeros(md.m)
    np.add.at(rows, md.index, md.value)
    assert np.allclose(md.row_upper, 0.4 * rows) and np.all(np.isneginf(md.row_lower))
    assert np.all(md.col_lower == 0) and np.all(md.col_upper == 1)
    fr = []
    for k in range(20):
        lo, up = smy.generators.fixings(6000 + k, md.col_lower, md.col_upper)
        fixed = lo == up
        fr.append(fixed.mean())
        assert np.all(np.isin(lo[fixed], (0.0, 1.0)))
        assert np.all(lo[~fixed] == 0) and np.all(up[~fixed] == 1)
        lo2, up2 = smy.generators.fixings(6000 + k, md.col_lower, md.col_upper)
        assert np.array_equal(lo, lo2) and np.array_equal(up, up2)
    assert 0.009 <= min(fr) and max(fr) <= 0.051 and len(set(fr)) > 5


def test_cpu_iterations_are_refused():
    lp, _, _ = O.make_problem(O.M.from_product(smy.generators.sparse_uniform(5, 10, 20, 3)))
    r = smy.solve_pdlp(lp.product(), smy.PdlpOptions(tol=1e-4, use_gpu=False))
    assert r.status == smy.NUMERICAL_ERROR and "GPU only" in r.message
