"""ctypes bindings of the native library (include/samaya_pdhg.h, include/samaya_pdlp.h).

The library is samaya_amd/_lib/libsamaya_amd.so, built in-tree by samaya_amd/build.py. There is
no fallback: load() raises when the library is missing, and every GPU entry point returns
SMY_NO_DEVICE without a gfx950 GPU.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libsamaya_amd.so"

SMY_OK, SMY_NO_DEVICE, SMY_TOO_LARGE, SMY_ERROR, SMY_INVALID, SMY_STATE = range(6)
SMY_FP64, SMY_FP32 = 0, 1
SMY_CURRENT, SMY_AVERAGE = 0, 1
SMY_PRECISION_FP64, SMY_PRECISION_MIXED = 0, 1
SMY_GEN_SPARSE_UNIFORM, SMY_GEN_SPARSE_POWERLAW, SMY_GEN_BLENDING, SMY_GEN_KNAPSACK = 1, 2, 3, 4

# Every symbol the headers declare (checked by the CPU test suite).
EXPORTS = (
    "smy_gpu_available", "smy_gpu_device_name", "smy_last_error", "smy_pdhg_create",
    "smy_pdhg_destroy", "smy_pdhg_set_original", "smy_pdhg_step", "smy_pdhg_advance",
    "smy_pdhg_restart", "smy_pdhg_download", "smy_pdhg_evaluate", "smy_pdhg_restart_distance",
    "smy_pdhg_restart_to", "smy_pdhg_evaluate_both", "smy_pdhg_run", "smy_pdhg_set_precision",
    "smy_pdhg_set_start", "smy_pdhg_last_run_ms", "smy_pdhg_time_kernels",
    "smy_pdhg_create_batch", "smy_pdhg_batch_size", "smy_pdhg_set_original_at",
    "smy_pdhg_set_start_at", "smy_pdhg_download_at", "smy_pdhg_run_batch",
    "smy_pdhg_evaluate_batch", "smy_pdhg_restart_distance_batch", "smy_pdhg_restart_to_batch",
    "smy_pdhg_set_precision_batch", "smy_pdhg_freeze", "smy_pdhg_create_shared",
    "smy_pdhg_detect_batch", "smy_pdhg_download_ray", "smy_pdhg_download_farkas", "smy_pdhg_run_adaptive",
    "smy_pdhg_get_tile", "smy_pdhg_set_tile", "smy_pdhg_get_bin_tile", "smy_pdhg_set_bin_tile",
    "smy_pdhg_set_fp64_accumulation", "smy_pdhg_time_adaptive_kernels", "smy_pdhg_inject_fault",
    "smy_pdhg_equilibrate", "smy_pdlp_equilibrate", "smy_lp_dual_bound",
    "smy_pdhg_set_halpern", "smy_pdhg_fixed_point_batch", "smy_pdhg_equilibrate_reference",
    "smy_pdhg_rule_init", "smy_pdhg_rule_decide", "smy_csr_spmv", "smy_csr_spmv_f32", "smy_csr_spmv_noalloc",
    "smy_spmv_threads_per_row", "smy_spmv_plan_create", "smy_spmv_plan_run",
    "smy_spmv_plan_destroy",
    "smy_model_create", "smy_model_generate", "smy_model_free", "smy_model_num_rows",
    "smy_model_num_cols", "smy_model_num_nonzeros", "smy_model_export", "smy_compute_scaling",
    "smy_make_problem", "smy_pdlp_params_default",
    "smy_solve_pdlp", "smy_solve_lp", "smy_solve_lp_batch", "smy_solve_lp_bounds_batch",
    "smy_generate_fixings", "smy_solve", "smy_verify_lp", "smy_verify_infeasibility",
    "smy_verify_unbounded_ray", "smy_gpu_info",
    "smy_read_mps", "smy_read_mps_text", "smy_model_name", "smy_model_names",
    "smy_model_set_names", "smy_model_set_integrality", "smy_model_integrality", "smy_write_mps",
    "smy_write_mps_text", "smy_write_solution", "smy_solution_text",
    "smy_mip_params_default", "smy_mip_solve", "smy_mip_propagate",
)


class SmyCsc(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("start", C.c_void_p),
                ("index", C.c_void_p), ("value", C.c_void_p)]


class SmyPdlpParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iterations", C.c_longlong), ("time_limit", C.c_double),
                ("use_gpu", C.c_int), ("precision", C.c_int), ("fp32_switch_tol", C.c_double),
                ("use_graph", C.c_int), ("detect", C.c_int), ("log_level", C.c_int),
                ("verify", C.c_int), ("polish_rounds", C.c_int), ("algorithm", C.c_int),
                ("share_matrix", C.c_int), ("device_detect", C.c_int), ("start_omega", C.c_double),
                ("adaptive_steps", C.c_int), ("fp32_stall_level", C.c_double),
                ("device_restart", C.c_int), ("fp64_accumulation", C.c_int),
                ("device_scaling", C.c_int), ("column_order", C.c_int),
                ("inject_fault_step", C.c_longlong), ("halpern", C.c_int),
                ("reflection", C.c_double), ("pid_primal_weight", C.c_int),
                ("farkas_projection", C.c_int)]


class SmyPdlpInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_longlong),
                ("fp32_iterations", C.c_longlong), ("primal_residual", C.c_double),
                ("dual_residual", C.c_double), ("gap", C.c_double),
                ("operator_norm", C.c_double), ("omega", C.c_double), ("setup_seconds", C.c_double),
                ("solve_seconds", C.c_double), ("kernel_ms", C.c_double), ("restarts", C.c_int),
                ("evaluations", C.c_int), ("message", C.c_char * 256),
                ("rejected_steps", C.c_longlong)]


class SmyPdhgAdaptive(C.Structure):
    _fields_ = [("eta", C.c_double), ("omega", C.c_double), ("weight", C.c_double),
                ("attempts", C.c_longlong), ("rejected", C.c_longlong)]


class SmyLpResult(C.Structure):
    _fields_ = [("status", C.c_int), ("verified", C.c_int), ("objective", C.c_double),
                ("max_primal_violation", C.c_double), ("max_dual_violation", C.c_double),
                ("iterations", C.c_longlong), ("fp32_iterations", C.c_longlong),
                ("rounds", C.c_int), ("pdlp", SmyPdlpInfo), ("message", C.c_char * 256),
                ("dual_bound", C.c_double)]


class SmyMipParams(C.Structure):
    _fields_ = [("lp", SmyPdlpParams), ("gap_tol", C.c_double), ("wave", C.c_int),
                ("max_nodes", C.c_longlong), ("time_limit", C.c_double),
                ("integrality_tol", C.c_double), ("feasibility_tol", C.c_double),
                ("log_level", C.c_int)]


class SmyMipResult(C.Structure):
    _fields_ = [("status", C.c_int), ("objective", C.c_double), ("bound", C.c_double),
                ("gap", C.c_double), ("nodes", C.c_longlong), ("waves", C.c_longlong),
                ("pruned_infeasible", C.c_longlong), ("pruned_bound", C.c_longlong),
                ("integral", C.c_longlong), ("unresolved", C.c_longlong), ("inexact", C.c_longlong),
                ("fixed_solves", C.c_longlong),
                ("lp_iterations", C.c_longlong), ("seconds", C.c_double),
                ("message", C.c_char * 256)]


_lib = None

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_longlong
f64 = C.c_double


def _proto(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)


def load() -> C.CDLL:
    """The native library; raises FileNotFoundError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python samaya_amd/build.py` "
            "(there is no CPU fallback for the PDLP iterations)")
    # PyTorch ships its own HIP runtime. When both are used in one process, torch must be loaded
    # first so that this library binds to the same runtime (libamdhip64 is shared by soname);
    # loading this library first leaves torch unable to see the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    _proto(lib, "smy_gpu_available", i32)
    _proto(lib, "smy_gpu_device_name", i32, C.c_char_p, C.c_size_t)
    _proto(lib, "smy_last_error", C.c_char_p)
    _proto(lib, "smy_pdhg_create", i32, C.POINTER(SmyCsc), C.POINTER(SmyCsc), vp, vp, vp, i32,
           C.POINTER(vp))
    _proto(lib, "smy_pdhg_destroy", None, vp)
    _proto(lib, "smy_pdhg_set_original", i32, vp, vp, vp, vp, vp, vp)
    _proto(lib, "smy_pdhg_step", i32, vp, f64, f64, f64, i64)
    _proto(lib, "smy_pdhg_advance", i32, vp)
    _proto(lib, "smy_pdhg_restart", i32, vp)
    _proto(lib, "smy_pdhg_download", i32, vp, vp, vp, vp, vp, vp, vp)
    _proto(lib, "smy_pdhg_evaluate", i32, vp, i32, vp)
    _proto(lib, "smy_pdhg_evaluate_both", i32, vp, vp)
    _proto(lib, "smy_pdhg_restart_distance", i32, vp, i32, C.POINTER(f64), C.POINTER(f64))
    _proto(lib, "smy_pdhg_restart_to", i32, vp, i32)
    _proto(lib, "smy_pdhg_run", i32, vp, f64, f64, f64, i64, i32)
    _proto(lib, "smy_pdhg_set_precision", i32, vp, i32)
    _proto(lib, "smy_pdhg_set_start", i32, vp, vp, vp)
    _proto(lib, "smy_pdhg_last_run_ms", i32, vp, C.POINTER(f64))
    _proto(lib, "smy_pdhg_time_kernels", i32, vp, i32, C.POINTER(f64), C.POINTER(f64))
    _proto(lib, "smy_pdhg_create_batch", i32, i32, C.POINTER(SmyCsc), C.POINTER(SmyCsc), vp, vp,
           vp, i32, C.POINTER(vp))
    _proto(lib, "smy_pdhg_batch_size", i32, vp)
    _proto(lib, "smy_pdhg_set_original_at", i32, vp, i32, vp, vp, vp, vp, vp)
    _proto(lib, "smy_pdhg_set_start_at", i32, vp, i32, vp, vp)
    _proto(lib, "smy_pdhg_download_at", i32, vp, i32, vp, vp, vp, vp, vp, vp)
    _proto(lib, "smy_pdhg_run_batch", i32, vp, vp, vp, f64, vp, i32)
    _proto(lib, "smy_pdhg_evaluate_batch", i32, vp, vp)
    _proto(lib, "smy_pdhg_restart_distance_batch", i32, vp, vp, vp)
    _proto(lib, "smy_pdhg_restart_to_batch", i32, vp, vp)
    _proto(lib, "smy_pdhg_set_precision_batch", i32, vp, vp)
    _proto(lib, "smy_pdhg_freeze", i32, vp, i32)
    _proto(lib, "smy_pdhg_detect_batch", i32, vp, vp, vp)
    _proto(lib, "smy_pdhg_download_ray", i32, vp, i32, vp)
    _proto(lib, "smy_pdhg_download_farkas", i32, vp, i32, vp)
    _proto(lib, "smy_pdhg_run_adaptive", i32, vp, vp, i32)
    _proto(lib, "smy_pdhg_set_halpern", i32, vp, i32, f64)
    _proto(lib, "smy_pdhg_fixed_point_batch", i32, vp, vp)
    _proto(lib, "smy_pdhg_get_tile", i32, vp, i32, i32, C.POINTER(i32), C.POINTER(i32))
    _proto(lib, "smy_pdhg_set_tile", i32, vp, i32, i32, i32, i32)
    _proto(lib, "smy_pdhg_get_bin_tile", i32, vp, i32, i32, i32, C.POINTER(i32), C.POINTER(i32),
           C.POINTER(i32))
    _proto(lib, "smy_pdhg_set_bin_tile", i32, vp, i32, i32, i32, i32, i32)
    _proto(lib, "smy_pdhg_set_fp64_accumulation", i32, vp, i32)
    _proto(lib, "smy_pdhg_inject_fault", i32, vp, C.c_longlong)
    _proto(lib, "smy_pdhg_equilibrate", i32, vp, i32, i32, vp, vp, vp)
    _proto(lib, "smy_pdhg_equilibrate_reference", i32, vp, i32, vp, vp, vp)
    _proto(lib, "smy_lp_dual_bound", i32, vp, vp, C.POINTER(C.c_double))
    _proto(lib, "smy_pdlp_equilibrate", i32, C.POINTER(SmyCsc), C.POINTER(SmyCsc), i32, i32, i32, vp, vp, vp)
    _proto(lib, "smy_pdhg_rule_init", i32, vp, vp)
    _proto(lib, "smy_pdhg_rule_decide", i32, vp, vp, C.c_longlong, vp)
    _proto(lib, "smy_pdhg_time_adaptive_kernels", i32, vp, i32, C.POINTER(f64), C.POINTER(f64),
           C.POINTER(f64))
    _proto(lib, "smy_pdhg_create_shared", i32, C.POINTER(SmyCsc), C.POINTER(SmyCsc), i32, vp, vp,
           vp, i32, C.POINTER(vp))
    _proto(lib, "smy_csr_spmv", i32, i32, vp, vp, vp, vp, vp, vp, i32)
    _proto(lib, "smy_csr_spmv_f32", i32, i32, vp, vp, vp, vp, vp, vp)
    _proto(lib, "smy_csr_spmv_noalloc", i32, i32, vp, vp, vp, vp, vp, i32, vp)
    _proto(lib, "smy_spmv_threads_per_row", i32, i64, i32)
    _proto(lib, "smy_spmv_plan_create", i32, i32, vp, vp, vp, vp, i32, C.POINTER(vp))
    _proto(lib, "smy_spmv_plan_run", i32, vp, vp, vp, vp)
    _proto(lib, "smy_spmv_plan_destroy", None, vp)
    _proto(lib, "smy_model_create", i32, i32, i32, i32, f64, vp, vp, vp, vp, vp, vp, vp, vp,
           C.POINTER(vp))
    _proto(lib, "smy_model_generate", i32, i32, C.c_uint64, i64, i64, i64, f64, C.POINTER(vp))
    _proto(lib, "smy_model_free", None, vp)
    _proto(lib, "smy_model_num_rows", i32, vp)
    _proto(lib, "smy_model_num_cols", i32, vp)
    _proto(lib, "smy_model_num_nonzeros", i64, vp)
    _proto(lib, "smy_model_export", i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(i32),
           C.POINTER(f64))
    _proto(lib, "smy_compute_scaling", i32, vp, vp, vp)
    _proto(lib, "smy_make_problem", i32, vp, vp, vp, vp, vp)
    _proto(lib, "smy_pdlp_params_default", None, C.POINTER(SmyPdlpParams))
    _proto(lib, "smy_solve_pdlp", i32, i32, i32, vp, vp, vp, vp, vp, vp,
           C.POINTER(SmyPdlpParams), vp, vp, C.POINTER(SmyPdlpInfo), vp, vp, vp, vp, vp, vp, vp,
           vp, i32, vp, vp, i32)
    _proto(lib, "smy_solve_lp", i32, vp, C.POINTER(SmyPdlpParams), C.POINTER(SmyLpResult), vp, vp,
           vp, vp, vp)
    _proto(lib, "smy_solve_lp_batch", i32, vp, i32, C.POINTER(SmyPdlpParams),
           C.POINTER(SmyLpResult), vp, vp, vp)
    _proto(lib, "smy_solve_lp_bounds_batch", i32, vp, i32, vp, vp, C.POINTER(SmyPdlpParams), vp,
           vp, C.POINTER(SmyLpResult), vp, vp, vp)
    _proto(lib, "smy_generate_fixings", i32, C.c_uint64, i32, f64, f64, vp, vp, vp, vp)
    _proto(lib, "smy_mip_params_default", None, C.POINTER(SmyMipParams))
    _proto(lib, "smy_mip_solve", i32, vp, vp, C.POINTER(SmyMipParams), C.POINTER(SmyMipResult), vp)
    _proto(lib, "smy_mip_propagate", i32, vp, vp, vp, vp, vp, vp, C.POINTER(i32))
    _proto(lib, "smy_solve", i32, vp, C.POINTER(f64))
    _proto(lib, "smy_verify_lp", i32, vp, vp, vp, f64, f64, vp, C.c_char_p, C.c_size_t)
    _proto(lib, "smy_verify_infeasibility", i32, vp, vp, C.c_char_p, C.c_size_t)
    _proto(lib, "smy_verify_unbounded_ray", i32, vp, vp, C.c_char_p, C.c_size_t)
    _proto(lib, "smy_gpu_info", i32, C.POINTER(i32), C.c_char_p, C.c_size_t)
    _proto(lib, "smy_read_mps", i32, C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_size_t)
    _proto(lib, "smy_read_mps_text", i32, C.c_char_p, C.c_size_t, i32, C.POINTER(vp), C.c_char_p,
           C.c_size_t, C.POINTER(i64))
    _proto(lib, "smy_model_name", i32, vp, C.c_char_p, C.c_size_t)
    _proto(lib, "smy_model_names", i64, vp, i32, vp, C.c_size_t)
    _proto(lib, "smy_model_set_names", i32, vp, i32, C.c_char_p, C.c_size_t)
    _proto(lib, "smy_model_set_integrality", i32, vp, vp)
    _proto(lib, "smy_model_integrality", i64, vp, vp)
    _proto(lib, "smy_write_mps", i32, vp, C.c_char_p, C.c_char_p, C.c_size_t)
    _proto(lib, "smy_write_mps_text", i64, vp, vp, C.c_size_t)
    _proto(lib, "smy_write_solution", i32, vp, C.c_char_p, C.POINTER(SmyLpResult), vp, vp, vp, vp,
           vp, C.c_char_p, C.c_size_t)
    _proto(lib, "smy_solution_text", i64, vp, C.POINTER(SmyLpResult), vp, vp, vp, vp, vp, vp,
           C.c_size_t)
    _lib = lib
    return lib


def last_error() -> str:
    return load().smy_last_error().decode(errors="replace")


class NativeError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} (code {code}): {last_error()}")
        self.code = code


def check(code: int, what: str) -> None:
    if code != SMY_OK:
        raise NativeError(code, what)
