"""The PDLP preconditioner on the host (smy_pdlp_equilibrate), pinned by what the mathematics
fixes: after infinity-norm Ruiz rounds every row and column of D_r A D_c has largest magnitude at
most 1 (each round maps a_ij to a_ij / sqrt(r_i c_j) <= sqrt(a_ij / r_i) <= 1) and, after ten, at
least 1 - 1e-2; the Pock-Chambolle step with alpha = 1 gives ||D_r A D_c||_2 <= 1 (Pock and
Chambolle, ICCV 2011, lemma 2); empty rows and columns keep factor 1; the reference's 2-norm
variant equals the oracle's (src/lp/pdlp.cpp:109-187). The device computes the same factors
(tests/test_gpu_scaling.py)."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O
from samaya_amd import _native as N

REFERENCE, ENHANCED = 0, 1


def csc_pair(A):
    A = sp.csc_matrix(A)
    A.sort_indices()
    T = A.T.tocsc()
    T.sort_indices()
    keep = []

    def view(M):
        start = M.indptr.astype(np.int64)
        index = M.indices.astype(np.int32)
        value = M.data.astype(np.float64)
        keep.extend([start, index, value])
        return N.SmyCsc(M.shape[0], M.shape[1], start.ctypes.data, index.ctypes.data, value.ctypes.data)
    return view(A), view(T), keep


def equilibrate(A, algorithm=ENHANCED, passes=10, pc=1):
    m, n = A.shape
    a, at, keep = csc_pair(A)
    rs, cs, out = np.zeros(m), np.zeros(n), np.zeros(3)
    rc = N.load().smy_pdlp_equilibrate(C.byref(a), C.byref(at), algorithm, passes, pc,
                                       cs.ctypes.data, rs.ctypes.data, out.ctypes.data)
    assert rc == 0
    return rs, cs, out


def random_matrices():
    rng = np.random.default_rng(17)
    for k in range(6):
        m, n = int(rng.integers(5, 60)), int(rng.integers(5, 80))
        A = rng.normal(size=(m, n)) * 10.0 ** rng.uniform(-3, 3, size=(m, n))
        A *= rng.random((m, n)) < 0.3
        if k % 2:
            A[rng.integers(m)] = 0.0        # an empty row
            A[:, rng.integers(n)] = 0.0     # an empty column
        yield A


def test_ruiz_rounds_bound_the_infinity_norms():
    for A in random_matrices():
        rs, cs, _ = equilibrate(A, passes=10, pc=0)
        S = np.abs(rs[:, None] * A * cs[None, :])
        rmax, cmax = S.max(axis=1), S.max(axis=0)
        rows, cols = np.abs(A).max(axis=1) > 0, np.abs(A).max(axis=0) > 0
        assert np.all(rmax <= 1 + 1e-15) and np.all(cmax <= 1 + 1e-15)
        assert np.all(rmax[rows] >= 1 - 5e-2) and np.all(cmax[cols] >= 1 - 5e-2)
        assert np.all(rs[~rows] == 1.0) and np.all(cs[~cols] == 1.0)
        rs2, cs2, _ = equilibrate(A, passes=2, pc=0)
        S2 = np.abs(rs2[:, None] * A * cs2[None, :])
        assert S2.max() <= 1 + 1e-15
        dev2 = max(1 - S2.max(axis=1)[rows].min(), 1 - S2.max(axis=0)[cols].min())
        dev10 = max(1 - rmax[rows].min(), 1 - cmax[cols].min())
        assert dev10 <= dev2


def test_pock_chambolle_bounds_the_operator_norm():
    for A in random_matrices():
        for passes in (0, 10):
            rs, cs, out = equilibrate(A, passes=passes, pc=1)
            S = rs[:, None] * A * cs[None, :]
            sigma = np.linalg.norm(S, 2)
            assert sigma <= 1 + 1e-12
            # The power-iteration bound: sqrt((1.01 s)^2 + 1) with s <= sigma_max(S) <= 1.
            est = np.sqrt(out[0] ** 2 - 1.0) / 1.01
            assert est <= sigma * (1 + 1e-12) and est >= 0.5 * sigma
            assert np.isclose(out[1], np.linalg.norm(A, axis=0).max(), rtol=1e-14)
            assert np.isclose(out[2], np.linalg.norm(A, axis=1).max(), rtol=1e-14)


def test_reference_variant_equals_the_oracle():
    for A in random_matrices():
        a, at, keep = csc_pair(A)
        m, n = A.shape
        S = sp.csc_matrix(A)
        S.sort_indices()
        lp = O.LP(m, n, S.indptr.astype(np.int64), S.indices.astype(np.int32), S.data.astype(float),
                  np.zeros(n + m), np.zeros(n + m), np.ones(n + m))
        want_rs, want_cs = O.oracle_ruiz(lp, 10)
        rs, cs, _ = equilibrate(A, algorithm=REFERENCE)
        assert np.array_equal(rs, want_rs) and np.array_equal(cs, want_cs)
