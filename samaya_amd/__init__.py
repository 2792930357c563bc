"""samaya_amd: samaya's PDLP LP-solver hot path rebuilt for AMD MI355X (gfx950).

A C++ host solver keeps PDLP's decisions; hand-written HIP kernels run the iterations. Python
here only marshals arrays into the native library's C ABI (include/samaya_pdhg.h,
include/samaya_pdlp.h, include/samaya_io.h, include/samaya_mip.h); it has no compute path of its
own.
"""
from ._native import LIB_PATH, load, last_error, NativeError
from .lp import (Model, NativeModel, LpProblem, compute_scaling, make_problem, PdlpOptions,
                 PdlpResult, LpResult, BatchTimes, solve_pdlp, solve_lp, solve_lp_batch,
                 solve_lp_bounds_batch, solve, verify_lp_optimality, verify_infeasibility, dual_bound,
                 verify_unbounded_ray, gpu_available, gpu_info, STATUS_NAMES, OPTIMAL,
                 INFEASIBLE, UNBOUNDED, ITERATION_LIMIT, TIME_LIMIT, NUMERICAL_ERROR)
from .mps import (MpsParseError, read_mps, read_mps_text, read_mps_native, read_mps_text_native,
                  mps_text, write_mps, solution_text, write_solution)
from .mip import MipOptions, MipResult, solve_mip, propagate_bounds
from . import generators

__all__ = [
    "LIB_PATH", "load", "last_error", "NativeError", "Model", "NativeModel", "LpProblem",
    "compute_scaling", "make_problem",
    "PdlpOptions", "PdlpResult", "LpResult", "BatchTimes", "solve_pdlp", "solve_lp",
    "solve_lp_batch", "solve_lp_bounds_batch", "solve",
    "verify_lp_optimality", "verify_infeasibility", "dual_bound", "verify_unbounded_ray", "gpu_available",
    "gpu_info", "STATUS_NAMES", "OPTIMAL", "INFEASIBLE", "UNBOUNDED", "ITERATION_LIMIT",
    "TIME_LIMIT", "NUMERICAL_ERROR", "generators", "MpsParseError", "read_mps", "read_mps_text",
    "read_mps_native", "read_mps_text_native", "mps_text", "write_mps", "solution_text",
    "write_solution", "MipOptions", "MipResult", "solve_mip", "propagate_bounds",
]
