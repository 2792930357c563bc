"""Branch and bound over batched PDLP node relaxations (include/samaya_mip.h, smy_mip_solve).

samaya's MIP search bounds each node by a warm-started dual simplex, one node at a time
(src/mip/branch_and_bound.hpp:133-145). Here the open nodes with the best bounds are solved as one
shared-matrix PDLP batch on the GPU, pruned by safe dual bounds and Farkas certificates; the
search itself (best bound, rounding for incumbents, most-fractional branching) runs in the native
library. This module only marshals the call."""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .lp import Model, NativeModel, PdlpOptions, STATUS_NAMES, _p


@dataclass
class MipOptions:
    lp: PdlpOptions = field(default_factory=PdlpOptions)  # the node relaxations
    gap_tol: float = 1e-4          # relative gap closing the search
    wave: int = 256                # open nodes bounded per batch
    max_nodes: int = 100000
    time_limit: float = math.inf
    integrality_tol: float = 1e-6
    feasibility_tol: float = 1e-6
    log_level: int = 0

    def native(self) -> N.SmyMipParams:
        p = N.SmyMipParams()
        N.load().smy_mip_params_default(C.byref(p))
        p.lp = self.lp.native()
        p.gap_tol = self.gap_tol
        p.wave = self.wave
        p.max_nodes = self.max_nodes
        p.time_limit = self.time_limit
        p.integrality_tol = self.integrality_tol
        p.feasibility_tol = self.feasibility_tol
        p.log_level = self.log_level
        return p


@dataclass
class MipResult:
    status: int
    objective: float     # the incumbent's (NaN without one), in the model's sense
    bound: float         # the best bound in the model's sense (maximization: upper)
    gap: float
    col_value: np.ndarray | None
    nodes: int
    waves: int
    pruned_infeasible: int
    pruned_bound: int
    integral: int
    unresolved: int
    inexact: int
    fixed_solves: int
    lp_iterations: int
    seconds: float
    message: str

    @property
    def status_name(self) -> str:
        return STATUS_NAMES[self.status]


def solve_mip(model: Model | NativeModel, integer=None, options: MipOptions | None = None) -> MipResult:
    """Branch and bound on `model` with integer columns `integer` (n booleans; None: the model's
    own integrality, e.g. from its MPS file)."""
    options = options or MipOptions()
    nm = model if isinstance(model, NativeModel) else model.native()
    flags = None if integer is None else np.ascontiguousarray(np.asarray(integer, dtype=bool), np.uint8)
    x = np.zeros(nm.n)
    res = N.SmyMipResult()
    params = options.native()
    rc = N.load().smy_mip_solve(nm.h, _p(flags), C.byref(params), C.byref(res), _p(x))
    N.check(rc, "mip_solve")
    has = not math.isnan(res.objective)
    return MipResult(res.status, res.objective, res.bound, res.gap, x if has else None, res.nodes,
                     res.waves, res.pruned_infeasible, res.pruned_bound, res.integral, res.unresolved,
                     res.inexact, res.fixed_solves, res.lp_iterations, res.seconds,
                     res.message.decode(errors="replace"))


def propagate_bounds(model: Model | NativeModel, integer=None):
    """(col_lower, col_upper, row_lower, row_upper, crossed): the bounds solve_mip starts from
    (smy_mip_propagate): column bounds the rows imply, and one-sided rows' activity ranges."""
    nm = model if isinstance(model, NativeModel) else model.native()
    flags = None if integer is None else np.ascontiguousarray(np.asarray(integer, dtype=bool), np.uint8)
    lo, up = np.zeros(nm.n), np.zeros(nm.n)
    rlo, rup = np.zeros(nm.m), np.zeros(nm.m)
    crossed = C.c_int()
    N.check(N.load().smy_mip_propagate(nm.h, _p(flags), _p(lo), _p(up), _p(rlo), _p(rup),
                                       C.byref(crossed)), "mip_propagate")
    return lo, up, rlo, rup, bool(crossed.value)
