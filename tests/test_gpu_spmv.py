"""CSR SpMV on the GPU against the oracle's brute-force product (tests/test_gpu.cpp:283-404,
699-756): the reference's 8 shapes, the empty-row + long-row edge case, the mixed 0 / 1-12 / 300
row lengths, rows longer than the LDS stage, bit-identical repeats, FP32 within its rounding
bound."""
import numpy as np
import pytest

from oracle_lib import oracle_csr_spmv

pytestmark = pytest.mark.gpu


def random_csr(rng, m, n, density):
    ptr, idx, val = [0], [], []
    for _ in range(m):
        cols = np.nonzero(rng.random(n) < density)[0]
        idx.extend(cols.tolist())
        val.extend(rng.uniform(-1, 1, len(cols)).tolist())
        ptr.append(len(idx))
    return np.array(ptr, np.int32), np.array(idx, np.int32), np.array(val)


def run_gpu(ptr, idx, val, x, fp32=False):
    import torch
    from samaya_amd.device import csr_spmv, csr_spmv_f32
    dev = "cuda:0"
    fdt = torch.float32 if fp32 else torch.float64
    p = torch.tensor(ptr, dtype=torch.int32, device=dev)
    i = torch.tensor(idx if len(idx) else np.zeros(1, np.int32), dtype=torch.int32, device=dev)
    v = torch.tensor(val if len(val) else np.zeros(1), dtype=fdt, device=dev)
    xt = torch.tensor(x, dtype=fdt, device=dev)
    y = torch.full((len(ptr) - 1,), float("nan"), dtype=fdt, device=dev)
    (csr_spmv_f32 if fp32 else csr_spmv)(p, i, v, xt, y)
    torch.cuda.synchronize()
    return y.cpu().numpy().astype(np.float64)


def rel_err(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)) if len(a) else 0.0


@pytest.mark.parametrize("m,n,density", [(10, 10, 0.5), (100, 50, 0.1), (1000, 500, 0.01),
                                          (100, 200, 0.05), (200, 100, 0.05), (1, 10, 1.0),
                                          (10, 1, 1.0), (100, 100, 0.0)])
def test_spmv_reference_shapes(gpu, m, n, density):
    rng = np.random.default_rng(123)
    ptr, idx, val = random_csr(rng, m, n, density)
    x = rng.uniform(-10, 10, n)
    y_cpu = oracle_csr_spmv(ptr, idx, val, x)
    y_gpu = run_gpu(ptr, idx, val, x)
    assert rel_err(y_gpu, y_cpu) <= 1e-12


def test_spmv_empty_rows_and_long_row(gpu):
    m, n = 100, 50
    ptr = np.zeros(m + 1, np.int32)
    ptr[2:] = n
    idx = np.arange(n, dtype=np.int32)
    val = np.ones(n)
    x = np.ones(n)
    y = run_gpu(ptr, idx, val, x)
    assert y[1] == n and np.all(y[[0] + list(range(2, m))] == 0.0)


def mixed_rows(seed=614, m=3000, n=2000, long_len=300):
    rng = np.random.default_rng(seed)
    ptr, idx, val = [0], [], []
    for i in range(m):
        ln = 0 if i % 97 == 0 else long_len if i % 41 == 0 else 1 + int(rng.integers(0, 12))
        idx.extend(rng.integers(0, n, ln).tolist())
        val.extend(rng.uniform(-1, 1, ln).tolist())
        ptr.append(len(idx))
    return np.array(ptr, np.int32), np.array(idx, np.int32), np.array(val), rng.uniform(-1, 1, n)


def test_spmv_mixed_lengths_repeatable(gpu):
    ptr, idx, val, x = mixed_rows()
    y_cpu = oracle_csr_spmv(ptr, idx, val, x)
    a = run_gpu(ptr, idx, val, x)
    b = run_gpu(ptr, idx, val, x)
    assert rel_err(a, y_cpu) <= 1e-12
    assert np.array_equal(a, b)


def test_spmv_rows_longer_than_lds_stage(gpu):
    # Rows of 1,025 ... 20,000 entries take the strided long-row path (stage = 1,024 entries).
    ptr, idx, val, x = mixed_rows(seed=7, m=500, n=50000, long_len=20000)
    ptr2 = ptr.copy()
    y_cpu = oracle_csr_spmv(ptr2, idx, val, x)
    y = run_gpu(ptr2, idx, val, x)
    assert rel_err(y, y_cpu) <= 1e-12
    assert np.array_equal(y, run_gpu(ptr2, idx, val, x))


def test_spmv_fp32_within_rounding_bound(gpu):
    ptr, idx, val, x = mixed_rows(seed=9)
    val32 = val.astype(np.float32).astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    y_exact = oracle_csr_spmv(ptr, idx, val32, x32)
    y = run_gpu(ptr, idx, val32, x32, fp32=True)
    # |error| <= gamma_k sum |a x| with k the row length, gamma_k = k u / (1 - k u), u = 2^-24.
    absval = oracle_csr_spmv(ptr, idx, np.abs(val32), np.abs(x32))
    k = np.diff(ptr).astype(np.float64) + 1
    u = 2.0 ** -24
    bound = (k * u / (1 - k * u)) * absval
    assert np.all(np.abs(y - y_exact) <= bound + 1e-30)
