"""Device-side infeasibility and unboundedness tests against a NumPy restatement of the
reference's host formulas (src/lp/pdlp.cpp:724-842 Farkas, :852-1042 improving ray), on device
states after real steps, in both batch layouts; and the solver with device detection against the
same solver detecting on the host (the reference's arithmetic) on the reference's infeasible and
unbounded models."""
import numpy as np
import pytest
import scipy.sparse as sp

import samaya_amd as smy
import samaya_amd.device as D
import samaya_amd.generators as G
from conftest import load_golden
from oracle_lib import M
from test_gpu_pdhg import equilibrated, identity_lp, transpose_csc

pytestmark = pytest.mark.gpu


def reference_numbers(lp, lower, upper, rs, cs, v, v_new, y_erg):
    """The 11 detection numbers by the reference's formulas (problem space before equilibration)."""
    n, m = lp.n, lp.m
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(m, n))
    yf = y_erg * rs
    y_max = np.max(np.abs(yf)) if m else 0.0
    w = np.concatenate([A.T @ yf, -yf])
    w_sum = np.sum(np.abs(w))
    c = np.where(np.abs(w) <= 1e-9 * w_sum, 0.0, w)
    fmin = fmax = 0.0
    nneg = npos = 0
    scale = 0.0
    for k in range(n + m):
        lo, up = lower[k], upper[k]
        lf, uf = np.isfinite(lo), np.isfinite(up)
        if c[k] > 0:
            fmin, nneg = (fmin + c[k] * lo, nneg) if lf else (fmin, nneg + 1)
            fmax, npos = (fmax + c[k] * up, npos) if uf else (fmax, npos + 1)
        elif c[k] < 0:
            fmin, nneg = (fmin + c[k] * up, nneg) if uf else (fmin, nneg + 1)
            fmax, npos = (fmax + c[k] * lo, npos) if lf else (fmax, npos + 1)
        t = (w[k] if k < n else -w[k]) / y_max if y_max > 0 else 0.0
        if abs(t) > 1e-9:
            if lf:
                scale = max(scale, abs(t * lo))
            if uf:
                scale = max(scale, abs(t * up))
    d = v_new - v
    d_max = np.max(np.abs(d))
    dn = d / d_max
    x = dn[:n] * cs
    r = dn[n:] / rs
    res = A @ x - r
    cost_dir = lp.cost[:n] @ x + lp.cost[n:] @ r
    fails = 0
    ray = np.zeros(n)
    fixed = lambda lo, up: up - lo <= 1e-10 * (1 + abs(lo))
    for j in range(n):
        lo, up = lower[j], upper[j]
        lb, ub = np.isfinite(lo), np.isfinite(up)
        xj = x[j]
        if lb and ub:
            fails += abs(xj) > (1e-3 if fixed(lo, up) else 1e-4)
            xj = 0.0
        elif lb:
            fails += xj < -1e-4
            xj = max(xj, 0.0)
        elif ub:
            fails += xj > 1e-4
            xj = min(xj, 0.0)
        ray[j] = xj
    av = A @ ray
    for i in range(m):
        lo, up = lower[n + i], upper[n + i]
        lb, ub = np.isfinite(lo), np.isfinite(up)
        if lb and ub:
            fails += abs(av[i]) > (1e-3 if fixed(lo, up) else 1e-4)
        elif lb:
            fails += av[i] < -1e-4
        elif ub:
            fails += av[i] > 1e-4
    return np.array([w_sum, y_max, fmin, fmax, nneg, npos, scale, d_max, res @ res, cost_dir, fails]), ray


def close(got, want, tol=1e-9):
    return np.all(np.abs(got - want) <= tol * (1.0 + np.abs(want)))


@pytest.mark.parametrize("layout", ["stacked", "shared"])
def test_detection_numbers_match_the_reference_formulas(gpu, layout):
    # Instances with free and one-sided columns and rows, so every branch of the tests is taken.
    md = G.sparse_uniform(81, 70, 160, 4)
    lp = identity_lp(M.from_product(md))
    n, m = lp.n, lp.m
    rng = np.random.default_rng(5)
    K = 3
    lows, ups = [], []
    for k in range(K):
        lo, up = lp.lower.copy(), lp.upper.copy()
        free = rng.random(n) < 0.2
        lo[:n][free] = -np.inf
        up[:n][rng.random(n) < 0.2] = np.inf
        if layout == "stacked":
            up[n:][rng.random(m) < 0.3] = np.inf
        lows.append(lo)
        ups.append(up)
    rs, cs = rng.uniform(0.5, 2.0, m), rng.uniform(0.5, 2.0, n)
    eqs = []
    for k in range(K):
        base = type(lp)(m, n, lp.start, lp.index, lp.value, lp.cost, lows[k], ups[k])
        eqs.append(equilibrated(base, rs, cs))
    a = (m, n, eqs[0].start, eqs[0].index, eqs[0].value)
    at = transpose_csc(eqs[0])
    if layout == "stacked":
        b = D.PdhgBatch.create([a] * K, [at] * K, [e.cost for e in eqs], [e.lower for e in eqs],
                               [e.upper for e in eqs])
    else:
        b = D.PdhgBatch.create_shared(a, at, eqs[0].cost, [e.lower for e in eqs], [e.upper for e in eqs])
    for k in range(K):
        b.set_original(k, cs, rs, lp.cost, lows[k], ups[k])
    b.run([0.3, 0.2, 0.25], [0.2, 0.35, 0.3], 1.0, [0] * K, 64)
    b.advance()
    b.run([0.3, 0.2, 0.25], [0.2, 0.35, 0.3], 1.0, [64] * K, 7)
    flags = [1, 0, 1]
    got = b.detect(flags)
    for k in (0, 2):
        st = b.download(k)
        want, ray = reference_numbers(lp, lows[k], ups[k], rs, cs, st["v"], st["v_new"], st["y_erg"])
        assert close(got[k], want), (k, got[k], want)
        assert close(b.download_ray(k), ray)
    # Flags 2 and 3: the Farkas numbers of the dual's sign projection (rows whose term of f_min,
    # resp. f_max, would be infinite set to 0); 5, 6, 7: the same for the dual's change since the
    # restart point; no ray test. The tested dual comes back with download_farkas.
    rst = [b.download(k)["y_erg"].copy() for k in range(K)]  # the averages become y_rst
    b.restart_to([1, -1, 1])
    b.advance()
    b.run([0.3, 0.2, 0.25], [0.2, 0.35, 0.3], 1.0, [0, 71, 0], 9)
    b.advance()
    b.run([0.3, 0.2, 0.25], [0.2, 0.35, 0.3], 1.0, [9, 80, 9], 5)
    for f in (2, 3, 5, 6, 7):
        got = b.detect([f, 0, f])
        for k in (0, 2):
            st = b.download(k)
            y = st["y_erg"].copy()
            if f & 4:
                y = y - rst[k]
            s = y if (f & 3) != 3 else -y
            if (f & 3) >= 2:
                cut = ((s < 0) & ~np.isfinite(lows[k][n:])) | ((s > 0) & ~np.isfinite(ups[k][n:]))
                y[cut] = 0.0
            want, _ = reference_numbers(lp, lows[k], ups[k], rs, cs, st["v"], st["v_new"], y)
            assert close(got[k][:7], want[:7]), (f, k, got[k], want)
            assert np.all(got[k][7:] == 0.0)
            assert close(b.download_farkas(k), y)
    for bad in (4, 8):
        with pytest.raises(RuntimeError):
            b.detect([bad, 0, 0])


def test_bounded_instances_skip_the_ray_test(gpu):
    # Every structural variable boxed: the LP has no improving ray; the device leaves the ray
    # slots 0 and computes the Farkas numbers as always.
    md = G.knapsack(82, 50, 120, 4)
    lp = identity_lp(M.from_product(md))
    n, m = lp.n, lp.m
    assert np.all(np.isfinite(lp.lower[:n])) and np.all(np.isfinite(lp.upper[:n]))
    rng = np.random.default_rng(6)
    rs, cs = rng.uniform(0.5, 2.0, m), rng.uniform(0.5, 2.0, n)
    eq = equilibrated(lp, rs, cs)
    a = (m, n, eq.start, eq.index, eq.value)
    b = D.PdhgBatch.create([a], [transpose_csc(eq)], [eq.cost], [eq.lower], [eq.upper])
    b.set_original(0, cs, rs, lp.cost, lp.lower, lp.upper)
    b.run([0.3], [0.3], 1.0, [0], 64)
    b.advance()
    b.run([0.3], [0.3], 1.0, [64], 3)
    got = b.detect([1])[0]
    st = b.download(0)
    want, _ = reference_numbers(lp, lp.lower, lp.upper, rs, cs, st["v"], st["v_new"], st["y_erg"])
    assert close(got[:7], want[:7])
    assert np.all(got[7:] == 0.0) and want[7] > 0.0


def test_device_and_host_detection_agree(gpu):
    fx = load_golden("pdlp_infeasible_unbounded")
    models = [M.from_json(c["model"]).product_model() for c in fx["infeasible"] + fx["unbounded"]]
    same_status = same_iters = 0
    for md in models:
        o = smy.PdlpOptions(tol=1e-4, max_iterations=100000)
        dev = smy.solve_lp(md, o)
        o.device_detect = False
        host = smy.solve_lp(md, o)
        same_status += dev.status == host.status
        same_iters += dev.iterations == host.iterations
    assert same_status == len(models)
    assert same_iters >= 0.9 * len(models)


def test_infeasible_nodes_are_found_in_a_shared_batch(gpu):
    # Knapsack nodes fixing heavy items to 1 overflow a row: infeasible; the rest are optimal.
    md = G.knapsack(82, 150, 400, 4)
    lows, ups = [], []
    A = sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))
    for k in range(8):
        lo, up = G.fixings(700 + k, md.col_lower, md.col_upper, 0.05, 0.15)
        lows.append(lo)
        ups.append(up)
    infeasible = [bool(np.any(A @ lo > md.row_upper + 1e-9)) for lo in lows]
    assert any(infeasible) and not all(infeasible)
    res, _ = smy.solve_lp_bounds_batch(md, lows, ups, smy.PdlpOptions(tol=1e-6))
    for r, inf in zip(res, infeasible):
        assert r.status == (smy.INFEASIBLE if inf else smy.OPTIMAL), (r.status, inf)
