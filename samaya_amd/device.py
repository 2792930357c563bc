"""Python mirror of samaya's GPU seam: gpu::PdhgDevice (src/gpu/pdhg_core.hpp:24-79) and
gpu::csr_spmv / spmv_threads_per_row (src/gpu/spmv.hpp:16-23), over the C ABI of
include/samaya_pdhg.h. Device arrays for csr_spmv are torch tensors on cuda:0 (memory plumbing
only; the product is the HIP kernel).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .lp import _f64, _p

CURRENT, AVERAGE = N.SMY_CURRENT, N.SMY_AVERAGE


def _csc(rows: int, cols: int, start: np.ndarray, index: np.ndarray, value: np.ndarray):
    keep = (np.ascontiguousarray(start, np.int64), np.ascontiguousarray(index, np.int32),
            _f64(value))
    s = N.SmyCsc(rows, cols, keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data)
    return s, keep


class PdhgDevice:
    """One LP's PDHG state on the GPU. create() returns None without a usable GPU or when the
    matrix is too large for 32-bit offsets (PdhgDevice::create's null); driver failures raise."""

    def __init__(self, h, n: int, m: int):
        self.lib = N.load()
        self.h = h
        self.n, self.m = n, m

    @classmethod
    def create(cls, A: tuple, At: tuple, cost, lower, upper, precision: str = "fp64"):
        """A = (rows, cols, start, index, value) CSC of the m x n matrix; At the same for A'."""
        lib = N.load()
        a, keep_a = _csc(*A)
        at, keep_at = _csc(*At)
        h = C.c_void_p()
        prec = {"fp64": N.SMY_FP64, "fp32": N.SMY_FP32}[precision]
        rc = lib.smy_pdhg_create(C.byref(a), C.byref(at), _p(_f64(cost)), _p(_f64(lower)),
                                 _p(_f64(upper)), prec, C.byref(h))
        if rc in (N.SMY_NO_DEVICE, N.SMY_TOO_LARGE):
            return None
        N.check(rc, "pdhg_create")
        return cls(h, A[1], A[0])

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.smy_pdhg_destroy(self.h)
            self.h = None

    def set_original(self, col_scale, row_scale, cost, lower, upper):
        N.check(self.lib.smy_pdhg_set_original(self.h, _p(_f64(col_scale)), _p(_f64(row_scale)),
                                               _p(_f64(cost)), _p(_f64(lower)), _p(_f64(upper))),
                "set_original")

    def step(self, tau: float, sigma: float, theta: float, iteration: int):
        N.check(self.lib.smy_pdhg_step(self.h, tau, sigma, theta, iteration), "step")

    def run(self, tau: float, sigma: float, theta: float, iter0: int, count: int):
        N.check(self.lib.smy_pdhg_run(self.h, tau, sigma, theta, iter0, count), "run")

    def last_run_ms(self) -> float:
        ms = C.c_double()
        N.check(self.lib.smy_pdhg_last_run_ms(self.h, C.byref(ms)), "last_run_ms")
        return ms.value

    def time_kernels(self, reps: int = 20) -> tuple[float, float]:
        """Average ms of the dual and primal step kernels (changes the iterates)."""
        d, p = C.c_double(), C.c_double()
        N.check(self.lib.smy_pdhg_time_kernels(self.h, reps, C.byref(d), C.byref(p)), "time_kernels")
        return d.value, p.value

    def advance(self):
        N.check(self.lib.smy_pdhg_advance(self.h), "advance")

    def time_adaptive_kernels(self, reps: int = 20) -> tuple[float, float, float]:
        """Average ms of the adaptive dual, primal and decision kernels (changes the state)."""
        d, p, q = C.c_double(), C.c_double(), C.c_double()
        N.check(self.lib.smy_pdhg_time_adaptive_kernels(self.h, reps, C.byref(d), C.byref(p), C.byref(q)),
                "time_adaptive_kernels")
        return d.value, p.value, q.value

    def set_fp64_accumulation(self, on: bool):
        """FP32 steps sum their products in FP64 (storage stays FP32)."""
        N.check(self.lib.smy_pdhg_set_fp64_accumulation(self.h, int(on)), "set_fp64_accumulation")

    def equilibrate(self, passes: int = 10, pock_chambolle: bool = True, algorithm: str = "enhanced"):
        """The preconditioner on the device (the enhanced one, or the reference's 2-norm variant),
        applied to this handle's (unscaled) LP: (col_scale, row_scale, (L, a_colmax, at_rowmax))."""
        cs, rs, out = np.zeros(self.n), np.zeros(self.m), np.zeros(3)
        if algorithm == "reference":
            rc = self.lib.smy_pdhg_equilibrate_reference(self.h, passes, _p(cs), _p(rs), _p(out))
        else:
            rc = self.lib.smy_pdhg_equilibrate(self.h, passes, int(pock_chambolle), _p(cs), _p(rs),
                                               _p(out))
        N.check(rc, "equilibrate")
        return cs, rs, tuple(out)

    def restart(self):
        N.check(self.lib.smy_pdhg_restart(self.h), "restart")

    def download(self):
        """dict of v, v_new, y, y_new, v_erg, y_erg (host copies)."""
        nt, m = self.n + self.m, self.m
        out = {k: np.zeros(nt) for k in ("v", "v_new", "v_erg")}
        out.update({k: np.zeros(m) for k in ("y", "y_new", "y_erg")})
        N.check(self.lib.smy_pdhg_download(self.h, _p(out["v"]), _p(out["v_new"]), _p(out["y"]),
                                           _p(out["y_new"]), _p(out["v_erg"]), _p(out["y_erg"])),
                "download")
        return out

    def evaluate(self, which: int) -> np.ndarray:
        out = np.zeros(7)
        N.check(self.lib.smy_pdhg_evaluate(self.h, which, _p(out)), "evaluate")
        return out

    def evaluate_both(self) -> np.ndarray:
        out = np.zeros(14)
        N.check(self.lib.smy_pdhg_evaluate_both(self.h, _p(out)), "evaluate_both")
        return out

    def restart_distance(self, which: int) -> tuple[float, float]:
        dv, dy = C.c_double(), C.c_double()
        N.check(self.lib.smy_pdhg_restart_distance(self.h, which, C.byref(dv), C.byref(dy)),
                "restart_distance")
        return dv.value, dy.value

    def restart_to(self, which: int):
        N.check(self.lib.smy_pdhg_restart_to(self.h, which), "restart_to")

    def set_precision(self, precision: str):
        N.check(self.lib.smy_pdhg_set_precision(
            self.h, {"fp64": N.SMY_FP64, "fp32": N.SMY_FP32}[precision]), "set_precision")

    def set_start(self, v, y):
        N.check(self.lib.smy_pdhg_set_start(self.h, _p(_f64(v)), _p(_f64(y))), "set_start")

    def set_halpern(self, on: bool, reflection: float = 1.0):
        """Restarted Halpern steps with this reflection (smy_pdhg_set_halpern)."""
        N.check(self.lib.smy_pdhg_set_halpern(self.h, int(on), reflection), "set_halpern")

    def fixed_point(self) -> tuple[float, float]:
        """||v+ - v||^2, ||y+ - y||^2 of the last Halpern step of the last run."""
        out = np.zeros(2)
        N.check(self.lib.smy_pdhg_fixed_point_batch(self.h, _p(out)), "fixed_point")
        return float(out[0]), float(out[1])


class PdhgBatch:
    """Many LPs' PDHG state on one device handle (smy_pdhg_create_batch): instance i is the LP of
    A[i] (m_i x n_i CSC), its transpose At[i] and cost / lower / upper[i] (n_i + m_i). Per-instance
    arguments are sequences of batch length; `which` entries are -1 (skip), CURRENT or AVERAGE."""

    _PREC = {"fp64": N.SMY_FP64, "fp32": N.SMY_FP32}

    def __init__(self, h, sizes):
        self.lib = N.load()
        self.h = h
        self.sizes = sizes  # (n_i, m_i)
        self.size = len(sizes)

    @classmethod
    def create(cls, A, At, cost, lower, upper, precision: str = "fp64"):
        lib = N.load()
        k = len(A)
        a = (N.SmyCsc * k)()
        at = (N.SmyCsc * k)()
        keep = []
        for i in range(k):
            a[i], ka = _csc(*A[i])
            at[i], kt = _csc(*At[i])
            keep += [ka, kt]
        vecs = [[_f64(v) for v in vs] for vs in (cost, lower, upper)]
        ptrs = [(C.c_void_p * k)(*[v.ctypes.data for v in vs]) for vs in vecs]
        h = C.c_void_p()
        rc = lib.smy_pdhg_create_batch(k, a, at, ptrs[0], ptrs[1], ptrs[2], cls._PREC[precision],
                                       C.byref(h))
        if rc in (N.SMY_NO_DEVICE, N.SMY_TOO_LARGE):
            return None
        N.check(rc, "pdhg_create_batch")
        return cls(h, [(A[i][1], A[i][0]) for i in range(k)])

    @classmethod
    def create_shared(cls, A: tuple, At: tuple, cost, lower, upper, precision: str = "fp64"):
        """One matrix and cost for every instance, instance i with bounds lower[i] / upper[i]
        (n + m; equal row parts): the shared-matrix batch (instance-minor iterates, SpMM)."""
        lib = N.load()
        a, keep_a = _csc(*A)
        at, keep_at = _csc(*At)
        k = len(lower)
        lo = [_f64(v) for v in lower]
        up = [_f64(v) for v in upper]
        lop = (C.c_void_p * k)(*[v.ctypes.data for v in lo])
        upp = (C.c_void_p * k)(*[v.ctypes.data for v in up])
        h = C.c_void_p()
        rc = lib.smy_pdhg_create_shared(C.byref(a), C.byref(at), k, _p(_f64(cost)), lop, upp,
                                        cls._PREC[precision], C.byref(h))
        if rc in (N.SMY_NO_DEVICE, N.SMY_TOO_LARGE):
            return None
        N.check(rc, "pdhg_create_shared")
        return cls(h, [(A[1], A[0])] * k)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.smy_pdhg_destroy(self.h)
            self.h = None

    def set_original(self, i, col_scale, row_scale, cost, lower, upper):
        N.check(self.lib.smy_pdhg_set_original_at(self.h, i, _p(_f64(col_scale)), _p(_f64(row_scale)),
                                                  _p(_f64(cost)), _p(_f64(lower)), _p(_f64(upper))),
                "set_original_at")

    def set_start(self, i, v, y):
        N.check(self.lib.smy_pdhg_set_start_at(self.h, i, _p(_f64(v)), _p(_f64(y))), "set_start_at")

    def run(self, tau, sigma, theta: float, iter0, count: int):
        t, s = _f64(tau), _f64(sigma)
        k0 = np.ascontiguousarray(iter0, np.int64)
        N.check(self.lib.smy_pdhg_run_batch(self.h, _p(t), _p(s), theta, _p(k0), count), "run_batch")

    def run_adaptive(self, state, count: int):
        """`count` attempts of the adaptive step-size rule. `state`: a ctypes array of
        SmyPdhgAdaptive (batch size), read and updated."""
        N.check(self.lib.smy_pdhg_run_adaptive(self.h, state, count), "run_adaptive")

    def tile(self, matrix: int, precision: str = "fp64"):
        """(per_wg, xcd) of A (matrix 0) or A' (1) for that value precision; None before tuning."""
        pw, x = C.c_int(), C.c_int()
        q = {"fp64": N.SMY_FP64, "fp32": N.SMY_FP32}[precision]
        rc = self.lib.smy_pdhg_get_tile(self.h, matrix, q, C.byref(pw), C.byref(x))
        if rc == N.SMY_STATE:
            return None
        N.check(rc, "get_tile")
        return pw.value, bool(x.value)

    def set_tile(self, matrix: int, precision: str, per_wg: int, xcd: bool):
        q = {"fp64": N.SMY_FP64, "fp32": N.SMY_FP32}[precision]
        N.check(self.lib.smy_pdhg_set_tile(self.h, matrix, q, per_wg, int(xcd)), "set_tile")

    def set_halpern(self, on: bool, reflection: float = 1.0):
        """Restarted Halpern steps with this reflection (smy_pdhg_set_halpern)."""
        N.check(self.lib.smy_pdhg_set_halpern(self.h, int(on), reflection), "set_halpern")

    def fixed_point(self) -> np.ndarray:
        """(K, 2): ||v+ - v||^2, ||y+ - y||^2 of each instance's last Halpern step."""
        out = np.zeros((self.size, 2))
        N.check(self.lib.smy_pdhg_fixed_point_batch(self.h, _p(out)), "fixed_point_batch")
        return out

    def set_fp64_accumulation(self, on: bool):
        """FP32 steps sum their products in FP64 (storage stays FP32)."""
        N.check(self.lib.smy_pdhg_set_fp64_accumulation(self.h, int(on)), "set_fp64_accumulation")

    def bin_tile(self, matrix: int, precision: str, size_bin: int):
        """(per_wg, xcd, split) of one size bin of A (0) or A' (1); None before tuning."""
        pw, x, sp = C.c_int(), C.c_int(), C.c_int()
        q = {"fp64": N.SMY_FP64, "fp32": N.SMY_FP32}[precision]
        rc = self.lib.smy_pdhg_get_bin_tile(self.h, matrix, q, size_bin, C.byref(pw), C.byref(x),
                                            C.byref(sp))
        if rc == N.SMY_STATE:
            return None
        N.check(rc, "get_bin_tile")
        return pw.value, bool(x.value), bool(sp.value)

    def set_bin_tile(self, matrix: int, precision: str, size_bin: int, per_wg: int, xcd: bool):
        """Fixes one size bin's tile; the bins are then launched apart."""
        q = {"fp64": N.SMY_FP64, "fp32": N.SMY_FP32}[precision]
        N.check(self.lib.smy_pdhg_set_bin_tile(self.h, matrix, q, size_bin, per_wg, int(xcd)),
                "set_bin_tile")

    @staticmethod
    def adaptive_state(eta, omega, weight=0.0, attempts=0, rejected=0):
        """A state array for run_adaptive (per-instance sequences or scalars)."""
        eta = np.atleast_1d(np.asarray(eta, float))
        k = len(eta)
        om = np.broadcast_to(np.asarray(omega, float), (k,))
        w = np.broadcast_to(np.asarray(weight, float), (k,))
        at = np.broadcast_to(np.asarray(attempts, np.int64), (k,))
        rj = np.broadcast_to(np.asarray(rejected, np.int64), (k,))
        arr = (N.SmyPdhgAdaptive * k)()
        for i in range(k):
            arr[i] = N.SmyPdhgAdaptive(float(eta[i]), float(om[i]), float(w[i]), int(at[i]), int(rj[i]))
        return arr

    def advance(self):
        N.check(self.lib.smy_pdhg_advance(self.h), "advance")

    def last_run_ms(self) -> float:
        ms = C.c_double()
        N.check(self.lib.smy_pdhg_last_run_ms(self.h, C.byref(ms)), "last_run_ms")
        return ms.value

    def time_kernels(self, reps: int = 20) -> tuple[float, float]:
        """Average ms of the dual and primal step kernels (changes the iterates)."""
        d, p = C.c_double(), C.c_double()
        N.check(self.lib.smy_pdhg_time_kernels(self.h, reps, C.byref(d), C.byref(p)), "time_kernels")
        return d.value, p.value

    def time_adaptive_kernels(self, reps: int = 20) -> tuple[float, float, float]:
        """Average ms of the adaptive dual, primal and decision kernels over every instance
        (stacked layout only; changes the state)."""
        d, p, q = C.c_double(), C.c_double(), C.c_double()
        N.check(self.lib.smy_pdhg_time_adaptive_kernels(self.h, reps, C.byref(d), C.byref(p), C.byref(q)),
                "time_adaptive_kernels")
        return d.value, p.value, q.value

    def evaluate(self) -> np.ndarray:
        """(batch, 14): the current iterate's 7 sums, then the average's."""
        out = np.zeros((self.size, 14))
        N.check(self.lib.smy_pdhg_evaluate_batch(self.h, _p(out)), "evaluate_batch")
        return out

    def restart_distance(self, which) -> np.ndarray:
        w = np.ascontiguousarray(which, np.int32)
        out = np.zeros((self.size, 2))
        N.check(self.lib.smy_pdhg_restart_distance_batch(self.h, _p(w), _p(out)),
                "restart_distance_batch")
        return out

    def restart_to(self, which):
        w = np.ascontiguousarray(which, np.int32)
        N.check(self.lib.smy_pdhg_restart_to_batch(self.h, _p(w)), "restart_to_batch")

    def set_precision(self, precision):
        """precision[i]: "fp64", "fp32" or -1 (unchanged)."""
        p = np.array([self._PREC[q] if isinstance(q, str) else -1 for q in precision], np.int32)
        N.check(self.lib.smy_pdhg_set_precision_batch(self.h, _p(p)), "set_precision_batch")

    def freeze(self, i: int):
        N.check(self.lib.smy_pdhg_freeze(self.h, i), "freeze")

    def detect(self, flags) -> np.ndarray:
        """(batch, 11): the device-side infeasibility / unboundedness numbers of the flagged
        instances (smy_pdhg_detect_batch)."""
        f = np.ascontiguousarray(flags, np.int32)
        out = np.zeros((self.size, 11))
        N.check(self.lib.smy_pdhg_detect_batch(self.h, _p(f), _p(out)), "detect_batch")
        return out

    def download_ray(self, i: int) -> np.ndarray:
        ray = np.zeros(self.sizes[i][0])
        N.check(self.lib.smy_pdhg_download_ray(self.h, i, _p(ray)), "download_ray")
        return ray

    def download_farkas(self, i: int) -> np.ndarray:
        """The dual the last detection with a flag above 1 tested for instance i."""
        y = np.zeros(self.sizes[i][1])
        N.check(self.lib.smy_pdhg_download_farkas(self.h, i, _p(y)), "download_farkas")
        return y

    def download(self, i: int):
        """dict of v, v_new, y, y_new, v_erg, y_erg of instance i (host copies)."""
        n, m = self.sizes[i]
        out = {k: np.zeros(n + m) for k in ("v", "v_new", "v_erg")}
        out.update({k: np.zeros(m) for k in ("y", "y_new", "y_erg")})
        N.check(self.lib.smy_pdhg_download_at(self.h, i, _p(out["v"]), _p(out["v_new"]), _p(out["y"]),
                                              _p(out["y_new"]), _p(out["v_erg"]), _p(out["y_erg"])),
                "download_at")
        return out


def spmv_threads_per_row(nnz: int, rows: int) -> int:
    return int(N.load().smy_spmv_threads_per_row(nnz, rows))


def csr_spmv(ptr, idx, val, x, y, threads_per_row: int = 32) -> None:
    """y = A x on the GPU for torch tensors on cuda:0 (int32 ptr/idx, float64 val/x/y)."""
    import torch
    for t, dt in ((ptr, torch.int32), (idx, torch.int32), (val, torch.float64),
                  (x, torch.float64), (y, torch.float64)):
        if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
            raise ValueError("csr_spmv expects contiguous cuda tensors (int32 ptr/idx, float64)")
    rows = ptr.numel() - 1
    stream = torch.cuda.current_stream().cuda_stream
    N.check(N.load().smy_csr_spmv(rows, ptr.data_ptr(), idx.data_ptr(), val.data_ptr(),
                                  x.data_ptr(), y.data_ptr(), stream, threads_per_row), "csr_spmv")


def csr_spmv_f32(ptr, idx, val, x, y) -> None:
    import torch
    rows = ptr.numel() - 1
    stream = torch.cuda.current_stream().cuda_stream
    N.check(N.load().smy_csr_spmv_f32(rows, ptr.data_ptr(), idx.data_ptr(), val.data_ptr(),
                                      x.data_ptr(), y.data_ptr(), stream), "csr_spmv_f32")


class SpmvPlan:
    """A row-block schedule over a CSR matrix already on the GPU (torch tensors)."""

    def __init__(self, ptr_host: np.ndarray, ptr, idx, val, precision: str = "fp64"):
        self.lib = N.load()
        self.keep = (np.ascontiguousarray(ptr_host, np.int32), ptr, idx, val)
        self.h = C.c_void_p()
        prec = {"fp64": N.SMY_FP64, "fp32": N.SMY_FP32}[precision]
        N.check(self.lib.smy_spmv_plan_create(len(ptr_host) - 1, _p(self.keep[0]), ptr.data_ptr(),
                                              idx.data_ptr(), val.data_ptr(), prec,
                                              C.byref(self.h)), "spmv_plan_create")

    def run(self, x, y, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        N.check(self.lib.smy_spmv_plan_run(self.h, x.data_ptr(), y.data_ptr(), s), "spmv_plan_run")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.smy_spmv_plan_destroy(self.h)
            self.h = None
