"""The PDHG device layer against the oracle, on the reference's own device-test inputs
(tests/test_gpu.cpp:514-673: seed 611 models with random equilibration scales U(0.5, 2), seed
612 for restarts; the models and scales are the golden fixtures gpu_evaluate / gpu_restart)."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import load_golden
from oracle_lib import LP, M, equilibrate_lp, oracle_pdhg_steps, oracle_sums

pytestmark = pytest.mark.gpu
CURRENT, AVERAGE = 0, 1


def identity_lp(md: M) -> LP:
    """make_problem(model, Scaling::identity) (tests/test_gpu.cpp:519)."""
    sense = -1.0 if md.maximize else 1.0
    cost = np.concatenate([sense * md.obj, np.zeros(md.m)])
    lower = np.concatenate([md.col_lower, md.row_lower])
    upper = np.concatenate([md.col_upper, md.row_upper])
    return LP(md.m, md.n, md.start.astype(np.int64), md.index.astype(np.int32),
              md.value.astype(np.float64), cost, lower, upper)


def equilibrated(lp: LP, rs, cs):
    return equilibrate_lp(lp, rs, cs)


def transpose_csc(lp: LP):
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n))
    T = A.T.tocsc()
    T.sort_indices()
    return (lp.n, lp.m, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data)


def make_device(lp_eq: LP, precision="fp64"):
    from samaya_amd.device import PdhgDevice
    return PdhgDevice.create((lp_eq.m, lp_eq.n, lp_eq.start, lp_eq.index, lp_eq.value),
                             transpose_csc(lp_eq), lp_eq.cost, lp_eq.lower, lp_eq.upper, precision)


def cases(name):
    for c in load_golden(name)["cases"]:
        md = M.from_json(c["model"])
        yield md, np.array(c["rs"]), np.array(c["cs"])


def close(a, b, tol):
    return np.all(np.abs(np.asarray(a) - np.asarray(b)) <= tol * (1.0 + np.abs(np.asarray(b))))


def test_device_steps_match_oracle_iterates(gpu):
    for md, rs, cs in cases("gpu_evaluate"):
        lp = identity_lp(md)
        eq = equilibrated(lp, rs, cs)
        dev = make_device(eq)
        for it in range(7):
            dev.step(0.3, 0.3, 1.0, it)
            dev.advance()
        dev.step(0.3, 0.3, 1.0, 7)
        got = dev.download()
        vn, yn, ve, ye = oracle_pdhg_steps(eq, 0.3, 0.3, 1.0, 8)
        # The device sums each row in another order than the host's column scatter; with these
        # fixed step sizes the iterates grow to ~1e8 in 8 steps, so the bar is the reference's
        # own for device-vs-host iterates (1e-10 relative, docs/GETTING_STARTED.md:246-248).
        assert close(got["v_new"], vn, 1e-10) and close(got["y_new"], yn, 1e-10)
        assert close(got["v_erg"], ve, 1e-10) and close(got["y_erg"], ye, 1e-10)


def test_evaluate_matches_host_sums_and_repeats(gpu):
    compared = 0
    for md, rs, cs in cases("gpu_evaluate"):
        lp = identity_lp(md)
        eq = equilibrated(lp, rs, cs)
        dev = make_device(eq)
        dev.set_original(cs, rs, lp.cost, lp.lower, lp.upper)
        for it in range(7):
            dev.step(0.3, 0.3, 1.0, it)
            dev.advance()
        dev.step(0.3, 0.3, 1.0, 7)
        got = dev.download()
        for which, v, y in ((CURRENT, got["v_new"], got["y_new"]), (AVERAGE, got["v_erg"], got["y_erg"])):
            d1 = dev.evaluate(which)
            d2 = dev.evaluate(which)
            host = oracle_sums(lp, rs, cs, v, y)
            assert close(d1, host, 1e-11), (d1, host)
            assert np.array_equal(d1, d2)
            compared += 1
        both = dev.evaluate_both()
        assert np.array_equal(both[:7], dev.evaluate(CURRENT))
        assert np.array_equal(both[7:], dev.evaluate(AVERAGE))
    assert compared == 40


def test_restart_distance_and_restart_to(gpu):
    md, rs, cs = next(cases("gpu_restart"))
    lp = identity_lp(md)
    eq = equilibrated(lp, rs, cs)
    dev = make_device(eq)
    for it in range(5):
        dev.step(0.3, 0.3, 1.0, it)
        dev.advance()
    dev.step(0.3, 0.3, 1.0, 5)
    got = dev.download()
    dv, dy = dev.restart_distance(AVERAGE)
    hv, hy = float(np.sum(got["v_erg"] ** 2)), float(np.sum(got["y_erg"] ** 2))
    assert abs(dv - hv) <= 1e-12 * (1 + hv) and abs(dy - hy) <= 1e-12 * (1 + hy)
    assert hv > 0
    dev.restart_to(AVERAGE)
    after = dev.download()
    for k in ("v", "v_new"):
        assert np.array_equal(after[k], got["v_erg"])
    for k in ("y", "y_new"):
        assert np.array_equal(after[k], got["y_erg"])
    assert dev.restart_distance(CURRENT) == (0.0, 0.0)


def test_graph_run_equals_single_steps(gpu):
    md, rs, cs = next(cases("gpu_evaluate"))
    eq = equilibrated(identity_lp(md), rs, cs)
    a, b = make_device(eq), make_device(eq)
    for rep in range(3):  # replays of the captured graph, with host work between them
        a.run(0.2, 0.4, 1.0, 64 * rep, 64)
        a.advance()
        for j in range(64):
            b.step(0.2, 0.4, 1.0, 64 * rep + j)
            b.advance()
        da, db = a.download(), b.download()
        for k in da:
            assert np.array_equal(da[k], db[k]), k


def test_restart_then_run_starts_without_momentum(gpu):
    md, rs, cs = next(cases("gpu_evaluate"))
    eq = equilibrated(identity_lp(md), rs, cs)
    dev = make_device(eq)
    for it in range(10):
        dev.step(0.3, 0.3, 1.0, it)
        dev.advance()
    dev.step(0.3, 0.3, 1.0, 10)
    dev.restart_to(CURRENT)
    dev.advance()
    start = dev.download()["v"]
    # From a restart point v = v_prev, a step with theta = 1 equals a step with theta = 0.
    t1 = make_device(eq)
    t1.set_start(start, dev.download()["y"])
    t1.step(0.3, 0.3, 0.0, 0)
    dev.step(0.3, 0.3, 1.0, 0)
    assert np.array_equal(dev.download()["v_new"], t1.download()["v_new"])


def test_fp32_steps_track_fp64(gpu):
    for md, rs, cs in list(cases("gpu_evaluate"))[:5]:
        eq = equilibrated(identity_lp(md), rs, cs)
        d64, d32 = make_device(eq), make_device(eq, "fp32")
        for it in range(20):
            for d in (d64, d32):
                d.step(0.3, 0.3, 1.0, it)
                d.advance()
        a, b = d64.download(), d32.download()
        scale = 1.0 + np.max(np.abs(a["v"]))
        assert np.max(np.abs(a["v"] - b["v"])) <= 1e-4 * scale
        # Switching precision converts the state: FP64 continuation from the FP32 point.
        d32.set_precision("fp64")
        c = d32.download()
        assert np.array_equal(c["v"], b["v"])
