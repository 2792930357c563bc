"""The GPU-failure path (the reference's src/lp/pdlp.cpp:228-246,341-360) through the test hook
smy_pdhg_inject_fault: a step entry point fails from the host side (no GPU fault), the solve
returns the error with its message, the handle is freed, and the next solve in the same process
runs normally and gives the bits of a solve that never saw a fault."""
import numpy as np
import pytest

import samaya_amd as smy
from samaya_amd import _native as N

pytestmark = pytest.mark.gpu

SMY_ERROR = 3


def small_lp():
    return smy.generators.sparse_uniform(7, 200, 500, 6)


def test_device_step_fails_at_the_injected_step_and_the_next_handle_runs(gpu):
    from samaya_amd.device import PdhgDevice
    import scipy.sparse as sp
    md = small_lp()
    lp = smy.make_problem(md)
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n))
    T = A.T.tocsc()
    T.sort_indices()
    at = (lp.n, lp.m, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data)

    def make():
        return PdhgDevice.create((lp.m, lp.n, lp.start, lp.index, lp.value), at, lp.cost,
                                 lp.lower, lp.upper, "fp64")

    dev = make()
    assert gpu.smy_pdhg_inject_fault(dev.h, 5) == 0
    for it in range(5):
        dev.step(0.1, 0.1, 1.0, it)
        dev.advance()
    rc = gpu.smy_pdhg_step(dev.h, 0.1, 0.1, 1.0, 5)
    assert rc == SMY_ERROR
    assert b"injected fault at step 5" in gpu.smy_last_error()
    rc = gpu.smy_pdhg_run(dev.h, 0.1, 0.1, 1.0, 5, 64)
    assert rc == SMY_ERROR
    got = dev.download()  # the handle stays usable for reads, then is freed
    del dev
    fresh = make()
    for it in range(5):
        fresh.step(0.1, 0.1, 1.0, it)
        fresh.advance()
    want = fresh.download()
    assert np.array_equal(got["v"], want["v"]) and np.array_equal(got["y"], want["y"])
    fresh.step(0.1, 0.1, 1.0, 5)  # no hook on this handle: the sixth step runs
    del fresh


def test_solve_reports_the_failure_then_the_next_solve_succeeds(gpu):
    md = small_lp()
    nm = md.native()
    clean = smy.solve_lp(nm, smy.PdlpOptions(tol=1e-6))
    assert clean.status == smy.OPTIMAL
    failed = smy.solve_lp(nm, smy.PdlpOptions(tol=1e-6, inject_fault_step=100))
    assert failed.status == smy.NUMERICAL_ERROR
    assert "the GPU failed during the solve" in failed.message
    assert "injected fault at step 100" in failed.message
    again = smy.solve_lp(nm, smy.PdlpOptions(tol=1e-6))
    assert again.status == smy.OPTIMAL and again.iterations == clean.iterations
    assert again.objective == clean.objective
    assert np.array_equal(again.col_value, clean.col_value)
    # Batches take the same path: every member reports the failure.
    res, _ = smy.solve_lp_batch([nm, nm], smy.PdlpOptions(tol=1e-6, inject_fault_step=64))
    assert all(r.status == smy.NUMERICAL_ERROR and "injected fault" in r.message for r in res)
    res, _ = smy.solve_lp_batch([nm, nm], smy.PdlpOptions(tol=1e-6))
    assert all(r.status == smy.OPTIMAL and r.objective == clean.objective for r in res)
