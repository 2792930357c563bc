"""The allocation-free SpMV entry point (smy_csr_spmv_noalloc) against the allocating one it is a
variant of (smy_csr_spmv / smy_csr_spmv_f32), through the C ABI: the same bits, on two calls in a
row (the second on a smaller matrix, so a stale schedule left in the shared workspace would show),
in FP64 and FP32; plus the size limit and the empty matrix."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

SMY_OK, SMY_TOO_LARGE, SMY_FP64, SMY_FP32 = 0, 2, 0, 1


def csr_case(seed, rows, cols):
    """Rows of mixed lengths: empty, short, 512/513 (around the schedule's long-row cut), 1,024 and
    3,000 (longer than a block), values N(0,1)."""
    rng = np.random.default_rng(seed)
    lens = rng.choice([0, 1, 7, 24, 512, 513, 1024, 3000], size=rows,
                      p=[0.1, 0.2, 0.3, 0.3, 0.03, 0.03, 0.02, 0.02])
    lens = np.minimum(lens, cols)
    indptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(cols, size=k, replace=False)) for k in lens]
                         ).astype(np.int32) if lens.sum() else np.zeros(0, np.int32)
    val = rng.normal(size=int(lens.sum()))
    return sp.csr_matrix((val, idx, indptr), shape=(rows, cols)), rng.normal(size=cols)


def run(lib, M, x, prec):
    import torch
    dt = torch.float64 if prec == SMY_FP64 else torch.float32
    ptr = torch.tensor(M.indptr.astype(np.int32), device="cuda")
    idx = torch.tensor(M.indices.astype(np.int32), device="cuda")
    val = torch.tensor(M.data, device="cuda", dtype=dt)
    xd = torch.tensor(x, device="cuda", dtype=dt)
    y_alloc = torch.full((M.shape[0],), np.nan, device="cuda", dtype=dt)
    y_free = torch.full((M.shape[0],), np.nan, device="cuda", dtype=dt)
    stream = torch.cuda.current_stream().cuda_stream
    rows = M.shape[0]
    if prec == SMY_FP64:
        rc = lib.smy_csr_spmv(rows, ptr.data_ptr(), idx.data_ptr(), val.data_ptr(), xd.data_ptr(),
                              y_alloc.data_ptr(), stream, 32)
    else:
        rc = lib.smy_csr_spmv_f32(rows, ptr.data_ptr(), idx.data_ptr(), val.data_ptr(),
                                  xd.data_ptr(), y_alloc.data_ptr(), stream)
    assert rc == SMY_OK, lib.smy_last_error()
    rc = lib.smy_csr_spmv_noalloc(rows, ptr.data_ptr(), idx.data_ptr(), val.data_ptr(),
                                  xd.data_ptr(), y_free.data_ptr(), prec, stream)
    assert rc == SMY_OK, lib.smy_last_error()
    torch.cuda.synchronize()
    return y_alloc.cpu().numpy(), y_free.cpu().numpy()


@pytest.mark.parametrize("prec", [SMY_FP64, SMY_FP32])
def test_noalloc_equals_allocating_entry_point_twice(gpu, prec):
    lib = gpu
    for seed, rows, cols in ((1, 60_000, 40_000), (2, 3_000, 5_000)):
        M, x = csr_case(seed, rows, cols)
        print(f"noalloc case seed {seed} rows {rows} nnz {M.nnz} prec {prec}", flush=True)
        y_alloc, y_free = run(lib, M, x, prec)
        assert np.array_equal(y_alloc, y_free)
        want = M @ (x if prec == SMY_FP64 else x.astype(np.float32).astype(np.float64))
        mag = abs(M) @ abs(x)
        tol = 1e-12 if prec == SMY_FP64 else 2e-4
        assert np.all(np.abs(y_free - want) <= tol * (1 + mag))


def test_noalloc_limits(gpu):
    lib = gpu
    null = C.c_void_p(0)
    assert lib.smy_csr_spmv_noalloc(0, null, null, null, null, null, SMY_FP64, null) == SMY_OK
    one = C.c_void_p(1)  # never dereferenced: the size check comes first
    rc = lib.smy_csr_spmv_noalloc((1 << 20) + 1, one, one, one, one, one, SMY_FP64, null)
    assert rc == SMY_TOO_LARGE
    assert b"1048576" in lib.smy_last_error()
