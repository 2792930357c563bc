"""The enhanced preconditioner on the device (smy_pdhg_equilibrate) against the host's
(smy_pdlp_equilibrate, pinned by the invariants of tests/test_scaling_cpu.py): the same factors
and norms bit for bit, the operator-norm bound to rounding, a handle scaled on the device steps
exactly like one given the host-scaled LP, and solves agree."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import samaya_amd as smy
from samaya_amd.device import PdhgDevice
from test_scaling_cpu import ENHANCED, REFERENCE, equilibrate as host_equilibrate

pytestmark = pytest.mark.gpu


def cases():
    rng = np.random.default_rng(23)
    for k in range(4):
        m, n = int(rng.integers(20, 300)), int(rng.integers(20, 500))
        A = rng.normal(size=(m, n)) * 10.0 ** rng.uniform(-2, 2, size=(m, n))
        A *= rng.random((m, n)) < 0.05
        A[rng.integers(m)] = 0.0
        A[:, rng.integers(n)] = 0.0
        yield sp.csc_matrix(A)
    md = smy.generators.c2(0.02)  # power-law columns up to 5,000 long
    yield sp.csc_matrix((md.value, md.index, md.start), shape=(md.m, md.n))


def device_of(A, cost, lower, upper, precision="fp64", T=None):
    A = sp.csc_matrix(A)
    A.sort_indices()
    if T is None:
        T = A.T.tocsc()
        T.sort_indices()
    m, n = A.shape
    return PdhgDevice.create((m, n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data),
                             (n, m, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data),
                             cost, lower, upper, precision)


def vectors(m, n, seed):
    rng = np.random.default_rng(seed)
    cost = np.concatenate([rng.uniform(-1, 1, n), np.zeros(m)])
    lower = np.concatenate([np.zeros(n), np.where(rng.random(m) < 0.5, -np.inf, -1.0)])
    upper = np.concatenate([np.full(n, 10.0), rng.uniform(0, 2, m)])
    return cost, lower, upper


def test_device_reference_factors_equal_the_hosts(gpu):
    # The reference's 2-norm variant (src/lp/pdlp.cpp:109-187) on the device: factors, norms and
    # the stop rule's pass count as the host's, bit for bit; the bound to rounding. The diagonal
    # cases run all 10 passes (the variant oscillates with period 2 there), the others stop early.
    mats = list(cases()) + [sp.csc_matrix(np.diag([4.0, 0.5])), sp.csc_matrix(np.diag([1.05, 0.97]))]
    for k, A in enumerate(mats):
        m, n = A.shape
        print(f"reference scaling case {k}: {m} x {n}, nnz {A.nnz}", flush=True)
        rs_h, cs_h, out_h = host_equilibrate(A, REFERENCE, 10, 0)
        dev = device_of(A, *vectors(m, n, k))
        cs_d, rs_d, out_d = dev.equilibrate(10, algorithm="reference")
        assert np.array_equal(cs_d, cs_h) and np.array_equal(rs_d, rs_h), k
        assert out_d[1] == out_h[1] and out_d[2] == out_h[2]
        assert abs(out_d[0] - out_h[0]) <= 1e-12 * out_h[0]
    # The closed forms of the period-2 oscillation: diag(4, 0.5) ends at all scales 1 after 10
    # passes, diag(1.05, 0.97) stops after one pass at 1 / a_ii.
    d = device_of(sp.csc_matrix(np.diag([4.0, 0.5])), *vectors(2, 2, 0)).equilibrate(10, algorithm="reference")
    assert np.array_equal(d[0], [1.0, 1.0]) and np.array_equal(d[1], [1.0, 1.0])
    d = device_of(sp.csc_matrix(np.diag([1.05, 0.97])), *vectors(2, 2, 0)).equilibrate(10, algorithm="reference")
    np.testing.assert_allclose(d[0], [1 / 1.05, 1 / 0.97], rtol=1e-15)


def test_device_factors_equal_the_hosts(gpu):
    for k, A in enumerate(cases()):
        m, n = A.shape
        print(f"scaling case {k}: {m} x {n}, nnz {A.nnz}", flush=True)
        rs_h, cs_h, out_h = host_equilibrate(A, ENHANCED, 10, 1)
        dev = device_of(A, *vectors(m, n, k))
        cs_d, rs_d, out_d = dev.equilibrate(10, True)
        assert np.array_equal(cs_d, cs_h) and np.array_equal(rs_d, rs_h)
        assert out_d[1] == out_h[1] and out_d[2] == out_h[2]
        assert abs(out_d[0] - out_h[0]) <= 1e-12 * out_h[0]


def test_device_scaled_handle_steps_like_a_host_scaled_one(gpu):
    for k, A in enumerate(cases()):
        m, n = A.shape
        cost, lower, upper = vectors(m, n, 100 + k)
        dev = device_of(A, cost, lower, upper)
        cs, rs, _ = dev.equilibrate(10, True)
        # The host-scaled LP: A_eq = (a rs_i) cs_j by columns; cost and bounds as Lane::prepare.
        A = sp.csc_matrix(A)
        A.sort_indices()
        rows = A.indices
        cols = np.repeat(np.arange(n), np.diff(A.indptr))
        Aq = sp.csc_matrix((A.data * rs[rows] * cs[cols], A.indices, A.indptr), shape=(m, n))
        # A' by columns scaled as the host scales it: (a cs_j) rs_i.
        T = A.T.tocsc()
        T.sort_indices()
        tcols = np.repeat(np.arange(m), np.diff(T.indptr))
        Tq = sp.csc_matrix((T.data * cs[T.indices] * rs[tcols], T.indices, T.indptr), shape=(n, m))
        c2, l2, u2 = cost.copy(), lower.copy(), upper.copy()
        c2[:n] *= cs
        l2[:n] /= cs
        u2[:n] /= cs
        c2[n:] *= rs
        l2[n:] *= rs
        u2[n:] *= rs
        ref = device_of(Aq, c2, l2, u2, T=Tq)
        for it in range(6):
            for d in (dev, ref):
                d.step(0.05, 0.05, 1.0, it)
                d.advance()
        got, want = dev.download(), ref.download()
        for key in ("v", "y", "v_erg", "y_erg"):
            assert np.array_equal(got[key], want[key]), (k, key)


def test_solves_with_device_and_host_scaling_agree(gpu):
    md = smy.generators.c2(0.01)
    nm = md.native()
    base = dict(tol=1e-6, precision="mixed", algorithm="enhanced", adaptive_steps=True)
    on = smy.solve_lp(nm, smy.PdlpOptions(**base))
    off = smy.solve_lp(nm, smy.PdlpOptions(**base, device_scaling=False))
    assert on.status == off.status == smy.OPTIMAL
    assert abs(on.objective - off.objective) <= 1e-6 * (1 + abs(off.objective))
    assert abs(on.iterations - off.iterations) <= 0.1 * off.iterations
