"""Host-side LP objects and solver entry points (Python mirror of samaya's interfaces).

  Model ............. include/samaya/model.hpp:41-80 (LP fields), CSC matrix A
  LpProblem ......... src/lp/simplex.hpp:18-26 (computational form, A m x n CSC)
  PdlpOptions ....... src/lp/pdlp.hpp:11-16 (+ precision, graph, detection switches)
  PdlpResult ........ src/lp/pdlp.hpp:19-30
  solve_pdlp ........ src/lp/pdlp.hpp:35
  solve_lp .......... solve_pdlp_lp, src/core/solver.cpp:392-477 (LP by PDLP through Solver::solve)
  verify_* .......... include/samaya/verify.hpp:42-67
  gpu_info .......... include/samaya/solver.hpp:93-97

All arithmetic happens in the native library (C++ host solver + HIP kernels); this module only
marshals numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

STATUS_NAMES = ("optimal", "infeasible", "unbounded", "iteration limit", "time limit",
                "numerical error")
OPTIMAL, INFEASIBLE, UNBOUNDED, ITERATION_LIMIT, TIME_LIMIT, NUMERICAL_ERROR = range(6)


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


@dataclass
class Model:
    """min/max obj'x + offset  s.t.  row_lower <= A x <= row_upper, col_lower <= x <= col_upper.

    A is given as CSC arrays (start int64[n+1], index int32[nnz] sorted within columns, value
    float64[nnz]); infinite bounds are +-inf.
    """

    m: int
    n: int
    start: np.ndarray
    index: np.ndarray
    value: np.ndarray
    obj: np.ndarray
    col_lower: np.ndarray
    col_upper: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    maximize: bool = False
    offset: float = 0.0
    # What an MPS file gives a model besides the LP (include/samaya/model.hpp name, row_names,
    # col_names, col_type): empty / None for models built from arrays.
    name: str = ""
    row_names: list | None = None
    col_names: list | None = None
    col_integer: np.ndarray | None = None  # uint8[n], 1 = integer (the LP path solves the relaxation)

    def __post_init__(self):
        self.start = np.ascontiguousarray(self.start, dtype=np.int64)
        self.index = np.ascontiguousarray(self.index, dtype=np.int32)
        self.value = _f64(self.value)
        for name in ("obj", "col_lower", "col_upper", "row_lower", "row_upper"):
            setattr(self, name, _f64(getattr(self, name)))

    @property
    def nnz(self) -> int:
        return int(self.start[-1])

    @classmethod
    def from_dense(cls, A, obj, col_lower, col_upper, row_lower, row_upper, maximize=False,
                   offset=0.0) -> "Model":
        A = np.asarray(A, dtype=np.float64)
        m, n = A.shape
        start, index, value = [0], [], []
        for j in range(n):
            rows = np.nonzero(A[:, j])[0]
            index.extend(rows.tolist())
            value.extend(A[rows, j].tolist())
            start.append(len(index))
        return cls(m, n, np.array(start), np.array(index, dtype=np.int32), np.array(value),
                   obj, col_lower, col_upper, row_lower, row_upper, maximize, offset)

    @classmethod
    def generate(cls, kind: int, seed: int, p0: int, p1: int, p2: int, p3: float = 0.0) -> "Model":
        """A synthetic instance from the library's deterministic generators."""
        lib = N.load()
        h = C.c_void_p()
        N.check(lib.smy_model_generate(kind, seed, p0, p1, p2, p3, C.byref(h)), "generate")
        try:
            return cls._export(h)
        finally:
            lib.smy_model_free(h)

    @classmethod
    def _export(cls, h) -> "Model":
        lib = N.load()
        m, n, nnz = lib.smy_model_num_rows(h), lib.smy_model_num_cols(h), lib.smy_model_num_nonzeros(h)
        start = np.empty(n + 1, np.int64)
        index = np.empty(nnz, np.int32)
        value = np.empty(nnz)
        obj, cl, cu = np.empty(n), np.empty(n), np.empty(n)
        rl, ru = np.empty(m), np.empty(m)
        mx = C.c_int()
        off = C.c_double()
        N.check(lib.smy_model_export(h, _p(start), _p(index), _p(value), _p(obj), _p(cl), _p(cu),
                                     _p(rl), _p(ru), C.byref(mx), C.byref(off)), "export")
        return cls(m, n, start, index, value, obj, cl, cu, rl, ru, bool(mx.value), off.value)

    def native(self) -> "NativeModel":
        return NativeModel(self)


class NativeModel:
    """A model copied into the native library (smy_model*)."""

    def __init__(self, md: Model):
        self.lib = N.load()
        self.h = C.c_void_p()
        N.check(self.lib.smy_model_create(md.m, md.n, int(md.maximize), md.offset, _p(md.start),
                                          _p(md.index), _p(md.value), _p(md.obj), _p(md.col_lower),
                                          _p(md.col_upper), _p(md.row_lower), _p(md.row_upper),
                                          C.byref(self.h)), "model_create")
        self.m, self.n = md.m, md.n
        for which, names in ((0, md.row_names), (1, md.col_names)):
            if names:
                blob = "".join(f"{x}\n" for x in names).encode()
                N.check(self.lib.smy_model_set_names(self.h, which, blob, len(blob)), "set_names")
        if md.col_integer is not None:
            marks = np.ascontiguousarray(md.col_integer, dtype=np.uint8)
            N.check(self.lib.smy_model_set_integrality(self.h, _p(marks)), "set_integrality")

    @classmethod
    def _adopt(cls, h, m: int, n: int) -> "NativeModel":
        """Wraps a handle the library made (it is freed with the object)."""
        self = cls.__new__(cls)
        self.lib, self.h, self.m, self.n = N.load(), h, m, n
        return self

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.smy_model_free(self.h)
            self.h = None


def compute_scaling(model: "Model | NativeModel"):
    """compute_scaling (src/linalg/scaling.cpp:130-200): (col_scale, row_scale), powers of two."""
    nm = model if isinstance(model, NativeModel) else model.native()
    cs, rs = np.zeros(nm.n), np.zeros(nm.m)
    N.check(N.load().smy_compute_scaling(nm.h, _p(cs), _p(rs)), "compute_scaling")
    return cs, rs


def make_problem(model: "Model") -> "LpProblem":
    """make_problem(model, compute_scaling(model.A)) (src/lp/lp_solver.cpp:28-51)."""
    nm = model.native()
    nt = model.n + model.m
    cost, lower, upper = np.zeros(nt), np.zeros(nt), np.zeros(nt)
    value = np.zeros(model.nnz)
    N.check(N.load().smy_make_problem(nm.h, _p(cost), _p(lower), _p(upper), _p(value)),
            "make_problem")
    return LpProblem(model.m, model.n, model.start.copy(), model.index.copy(), value, cost, lower,
                     upper)


@dataclass
class LpProblem:
    """Computational form: min cost'v s.t. [A -I] v = 0, lower <= v <= upper (v = (x, r))."""

    m: int
    n: int
    start: np.ndarray
    index: np.ndarray
    value: np.ndarray
    cost: np.ndarray
    lower: np.ndarray
    upper: np.ndarray

    def __post_init__(self):
        self.start = np.ascontiguousarray(self.start, dtype=np.int64)
        self.index = np.ascontiguousarray(self.index, dtype=np.int32)
        self.value = _f64(self.value)
        self.cost, self.lower, self.upper = _f64(self.cost), _f64(self.lower), _f64(self.upper)


@dataclass
class PdlpOptions:
    tol: float = 1e-8
    max_iterations: int = -1
    time_limit: float = math.inf
    use_gpu: bool = True
    precision: str = "fp64"          # "fp64" or "mixed"
    fp32_switch_tol: float = 1e-5
    use_graph: bool = True
    detect: bool = True
    log_level: int = 0
    verify: bool = True
    polish_rounds: int = 0
    algorithm: str = "reference"    # "reference" or "enhanced"
    share_matrix: bool = True       # batches of one model: the shared-matrix (SpMM) batch
    device_detect: bool = True      # infeasibility / unboundedness tests on the GPU (else host)
    start_omega: float = 0.0        # primal weight to start with (<= 0: 1), with a warm start
    adaptive_steps: bool = False    # PDLP's adaptive step sizes (decided on the GPU)
    fp32_stall_level: float = 1e-4  # mixed: FP32 plateaus below this switch to FP64 (0: never)
    device_restart: bool = True     # restart decisions, distances and primal weights on the GPU
    fp64_accumulation: bool = False  # mixed: FP32 steps sum their products in FP64
    device_scaling: bool = True      # one LP: preconditioner and norm bound on the GPU
    column_order: str = "auto"       # model-level solves: "auto", "natural", "degree" or "touch"
    inject_fault_step: int = -1      # test hook: the device fails its steps after this many
    halpern: bool = False            # restarted Halpern PDHG with reflection (no averages)
    reflection: float = 1.0          # Halpern steps: the reflection in [0, 1]
    pid_primal_weight: bool = False  # primal weight by a PI controller on log(dy / dx)
    farkas_projection: int = -1      # sign-projected Farkas test too (-1: unless algorithm reference)

    def native(self) -> N.SmyPdlpParams:
        p = N.SmyPdlpParams()
        N.load().smy_pdlp_params_default(C.byref(p))
        p.tol = self.tol
        p.max_iterations = self.max_iterations
        p.time_limit = self.time_limit
        p.use_gpu = int(self.use_gpu)
        p.precision = {"fp64": N.SMY_PRECISION_FP64, "mixed": N.SMY_PRECISION_MIXED}[self.precision]
        p.fp32_switch_tol = self.fp32_switch_tol
        p.use_graph = int(self.use_graph)
        p.detect = int(self.detect)
        p.log_level = self.log_level
        p.verify = int(self.verify)
        p.polish_rounds = self.polish_rounds
        p.algorithm = {"reference": 0, "enhanced": 1}[self.algorithm]
        p.share_matrix = int(self.share_matrix)
        p.device_detect = int(self.device_detect)
        p.start_omega = self.start_omega
        p.adaptive_steps = int(self.adaptive_steps)
        p.fp32_stall_level = self.fp32_stall_level
        p.device_restart = int(self.device_restart)
        p.fp64_accumulation = int(self.fp64_accumulation)
        p.inject_fault_step = int(self.inject_fault_step)
        p.device_scaling = int(self.device_scaling)
        p.column_order = {"auto": 0, "natural": 1, "degree": 2, "touch": 3}[self.column_order]
        p.halpern = int(self.halpern)
        p.reflection = self.reflection
        p.pid_primal_weight = int(self.pid_primal_weight)
        p.farkas_projection = int(self.farkas_projection)
        return p


@dataclass
class PdlpResult:
    status: int
    x: np.ndarray
    y: np.ndarray
    dual_ray: np.ndarray | None
    primal_ray: np.ndarray | None
    iterations: int
    fp32_iterations: int
    primal_residual: float
    dual_residual: float
    gap: float
    operator_norm: float
    setup_seconds: float
    solve_seconds: float
    kernel_ms: float
    message: str
    restarts: list = field(default_factory=list)   # (iter, to_average, kkt, omega)
    evals: list = field(default_factory=list)      # (iter, (cur p, d, g), (avg p, d, g))
    omega: float = 1.0                             # the final primal weight
    rejected_steps: int = 0                        # adaptive steps: rejected attempts

    @property
    def status_name(self) -> str:
        return STATUS_NAMES[self.status]


def solve_pdlp(lp: LpProblem, options: PdlpOptions | None = None, start_x=None,
               start_y=None) -> PdlpResult:
    """samaya::solve_pdlp (src/lp/pdlp.hpp:35) on the GPU."""
    options = options or PdlpOptions()
    lib = N.load()
    m, n = lp.m, lp.n
    x, y = np.zeros(n + m), np.zeros(m)
    dray, pray = np.zeros(m), np.zeros(n)
    cap = 1 << 16
    r_it = np.zeros(cap, np.int64)
    r_av = np.zeros(cap, np.int32)
    r_kkt, r_om = np.zeros(cap), np.zeros(cap)
    ecap = 1 << 16
    e_it = np.zeros(ecap, np.int64)
    e_ms = np.zeros(6 * ecap)
    info = N.SmyPdlpInfo()
    params = options.native()
    sx = None if start_x is None else _f64(start_x)
    sy = None if start_y is None else _f64(start_y)
    lib.smy_solve_pdlp(m, n, _p(lp.start), _p(lp.index), _p(lp.value), _p(lp.cost), _p(lp.lower),
                       _p(lp.upper), C.byref(params), _p(sx), _p(sy), C.byref(info), _p(x), _p(y),
                       _p(dray), _p(pray), _p(r_it), _p(r_av), _p(r_kkt), _p(r_om), cap, _p(e_it),
                       _p(e_ms), ecap)
    nr = min(info.restarts, cap)
    ne = min(info.evaluations, ecap)
    return PdlpResult(
        status=info.status, x=x, y=y,
        dual_ray=dray if info.status == INFEASIBLE else None,
        primal_ray=pray if info.status == UNBOUNDED else None,
        iterations=info.iterations, fp32_iterations=info.fp32_iterations,
        primal_residual=info.primal_residual, dual_residual=info.dual_residual, gap=info.gap,
        operator_norm=info.operator_norm, omega=info.omega, rejected_steps=info.rejected_steps,
        setup_seconds=info.setup_seconds,
        solve_seconds=info.solve_seconds, kernel_ms=info.kernel_ms,
        message=info.message.decode(errors="replace"),
        restarts=[(int(r_it[k]), int(r_av[k]), float(r_kkt[k]), float(r_om[k])) for k in range(nr)],
        evals=[(int(e_it[k]), tuple(e_ms[6 * k:6 * k + 3]), tuple(e_ms[6 * k + 3:6 * k + 6]))
               for k in range(ne)],
    )


@dataclass
class LpResult:
    status: int
    verified: bool
    objective: float
    col_value: np.ndarray
    row_dual: np.ndarray
    col_dual: np.ndarray
    row_activity: np.ndarray
    certificate: np.ndarray | None
    max_primal_violation: float
    max_dual_violation: float
    iterations: int
    fp32_iterations: int
    rounds: int
    primal_residual: float
    dual_residual: float
    gap: float
    setup_seconds: float
    solve_seconds: float
    kernel_ms: float
    message: str
    omega: float = 1.0  # the last solve's final primal weight (start_omega of a warm start)
    rejected_steps: int = 0  # adaptive steps: rejected attempts of the last solve
    dual_bound: float = float("nan")  # optimal answers: weak-duality bound from the row duals

    @property
    def status_name(self) -> str:
        return STATUS_NAMES[self.status]


def solve_lp(model: Model | NativeModel, options: PdlpOptions | None = None) -> LpResult:
    """LP by PDLP on the GPU with verification (solve_pdlp_lp, src/core/solver.cpp:392-477)."""
    options = options or PdlpOptions()
    nm = model if isinstance(model, NativeModel) else model.native()
    lib = N.load()
    m, n = nm.m, nm.n
    x, yd, d, act = np.zeros(n), np.zeros(m), np.zeros(n), np.zeros(m)
    cert = np.zeros(max(m, n))
    res = N.SmyLpResult()
    params = options.native()
    lib.smy_solve_lp(nm.h, C.byref(params), C.byref(res), _p(x), _p(yd), _p(d), _p(act), _p(cert))
    return _lp_result(res, x, yd, d, act, cert)


def _lp_result(res: "N.SmyLpResult", x, yd, d, act, cert) -> LpResult:
    st = res.status
    m, n = len(yd), len(x)
    return LpResult(
        status=st, verified=bool(res.verified), objective=res.objective, col_value=x,
        row_dual=yd, col_dual=d, row_activity=act,
        certificate=(None if cert is None else cert[:m] if st == INFEASIBLE
                     else cert[:n] if st == UNBOUNDED else None),
        max_primal_violation=res.max_primal_violation, max_dual_violation=res.max_dual_violation,
        iterations=res.iterations, fp32_iterations=res.fp32_iterations, rounds=res.rounds,
        primal_residual=res.pdlp.primal_residual, dual_residual=res.pdlp.dual_residual,
        gap=res.pdlp.gap, setup_seconds=res.pdlp.setup_seconds,
        solve_seconds=res.pdlp.solve_seconds, kernel_ms=res.pdlp.kernel_ms,
        message=res.message.decode(errors="replace"), omega=res.pdlp.omega,
        rejected_steps=res.pdlp.rejected_steps, dual_bound=res.dual_bound)


@dataclass
class BatchTimes:
    setup_seconds: float   # host setup of every LP and the upload, summed over rounds
    solve_seconds: float   # setup included, summed over rounds
    kernel_ms: float       # device time of the step runs


def solve_lp_batch(models, options: PdlpOptions | None = None
                   ) -> tuple[list[LpResult], BatchTimes]:
    """solve_lp of every model, their PDLP iterations batched on one device handle: each result
    equals solve_lp of that model alone (no reference counterpart; the reference solves one LP at
    a time). col_dual, row_activity and certificates are not returned by the batch."""
    options = options or PdlpOptions()
    nms = [m if isinstance(m, NativeModel) else m.native() for m in models]
    lib = N.load()
    k = len(nms)
    xs = [np.zeros(nm.n) for nm in nms]
    ys = [np.zeros(nm.m) for nm in nms]
    hs = (C.c_void_p * k)(*[nm.h for nm in nms])
    xp = (C.c_void_p * k)(*[x.ctypes.data for x in xs])
    yp = (C.c_void_p * k)(*[y.ctypes.data for y in ys])
    res = (N.SmyLpResult * k)()
    times = np.zeros(3)
    params = options.native()
    rc = lib.smy_solve_lp_batch(hs, k, C.byref(params), res, xp, yp, _p(times))
    N.check(rc, "solve_lp_batch")
    out = [_lp_result(res[i], xs[i], ys[i], None, None, None) for i in range(k)]
    return out, BatchTimes(float(times[0]), float(times[1]), float(times[2]))


def solve_lp_bounds_batch(model, col_lower, col_upper, options: PdlpOptions | None = None,
                          start=None) -> tuple[list[LpResult], BatchTimes]:
    """Node relaxations of one model: solve_lp with column bounds col_lower[k] / col_upper[k] for
    every k, batched on one shared matrix (no reference counterpart). `start` = (column values,
    row duals) of a model-space point (e.g. the root LP's answer) warm-starts every relaxation."""
    options = options or PdlpOptions()
    nm = model if isinstance(model, NativeModel) else model.native()
    lib = N.load()
    k = len(col_lower)
    lo = [_f64(v) for v in col_lower]
    up = [_f64(v) for v in col_upper]
    lop = (C.c_void_p * k)(*[v.ctypes.data for v in lo])
    upp = (C.c_void_p * k)(*[v.ctypes.data for v in up])
    xs = [np.zeros(nm.n) for _ in range(k)]
    ys = [np.zeros(nm.m) for _ in range(k)]
    xp = (C.c_void_p * k)(*[x.ctypes.data for x in xs])
    yp = (C.c_void_p * k)(*[y.ctypes.data for y in ys])
    sx = sy = None
    if start is not None:
        sx, sy = _f64(start[0]), _f64(start[1])
    res = (N.SmyLpResult * k)()
    times = np.zeros(3)
    params = options.native()
    rc = lib.smy_solve_lp_bounds_batch(nm.h, k, lop, upp, C.byref(params), _p(sx), _p(sy), res, xp,
                                       yp, _p(times))
    N.check(rc, "solve_lp_bounds_batch")
    out = [_lp_result(res[i], xs[i], ys[i], None, None, None) for i in range(k)]
    return out, BatchTimes(float(times[0]), float(times[1]), float(times[2]))


def verify_lp_optimality(model: Model | NativeModel, x, y, primal_tol=1e-6, dual_tol=1e-6):
    """(ok, max primal violation, max dual violation, objective, message)."""
    nm = model if isinstance(model, NativeModel) else model.native()
    out = np.zeros(3)
    msg = C.create_string_buffer(256)
    ok = N.load().smy_verify_lp(nm.h, _p(_f64(x)), _p(_f64(y)), primal_tol, dual_tol, _p(out), msg, 256)
    return bool(ok == 1), float(out[0]), float(out[1]), float(out[2]), msg.value.decode()


def dual_bound(model: Model | NativeModel, y) -> float:
    """Weak-duality bound on the model's optimal value from row duals y (d = c - A'y convention):
    lower for minimization, upper for maximization (smy_lp_dual_bound)."""
    nm = model if isinstance(model, NativeModel) else model.native()
    out = C.c_double()
    N.check(N.load().smy_lp_dual_bound(nm.h, _p(_f64(y)), C.byref(out)), "lp_dual_bound")
    return out.value


def verify_infeasibility(model: Model | NativeModel, y):
    nm = model if isinstance(model, NativeModel) else model.native()
    msg = C.create_string_buffer(256)
    ok = N.load().smy_verify_infeasibility(nm.h, _p(_f64(y)), msg, 256)
    return bool(ok == 1), msg.value.decode()


def verify_unbounded_ray(model: Model | NativeModel, ray):
    nm = model if isinstance(model, NativeModel) else model.native()
    msg = C.create_string_buffer(256)
    ok = N.load().smy_verify_unbounded_ray(nm.h, _p(_f64(ray)), msg, 256)
    return bool(ok == 1), msg.value.decode()


def gpu_available() -> bool:
    return bool(N.load().smy_gpu_available())


def gpu_info() -> tuple[bool, str]:
    """GpuInfo: (built with GPU kernels, device name or '')."""
    built = C.c_int()
    buf = C.create_string_buffer(256)
    N.load().smy_gpu_info(C.byref(built), buf, 256)
    return bool(built.value), buf.value.decode()


def solve(model: Model | NativeModel) -> tuple[int, float]:
    """samaya_solve (include/samaya_c.h:41): (samaya Status code, objective)."""
    nm = model if isinstance(model, NativeModel) else model.native()
    obj = C.c_double()
    st = N.load().smy_solve(nm.h, C.byref(obj))
    return st, obj.value
