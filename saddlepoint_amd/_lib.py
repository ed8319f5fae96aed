"""Loads ``libsaddlepoint_hip.so`` (the C ABI of include/saddlepoint_hip.h).

The library and its gfx950 code object are built in-tree into
``saddlepoint_amd/lib/`` by ``make -C saddlepoint_amd/csrc``. There is no
fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"
LIB_PATH = LIB_DIR / "libsaddlepoint_hip.so"
CODE_OBJECT = LIB_DIR / "saddlepoint_kernels.hsaco"

SP_OK = 0
ERRORS = {1: "UNAVAILABLE", 2: "HIP", 3: "SHAPE", 4: "TOO_LARGE", 5: "INVALID", 6: "PROBLEM", 7: "NUMERICAL",
          8: "MEMORY"}
SP_ERR_UNAVAILABLE = 1
SP_ERR_HIP = 2
SP_ERR_SHAPE = 3
SP_ERR_TOO_LARGE = 4
SP_ERR_INVALID = 5
SP_ERR_PROBLEM = 6
SP_ERR_NUMERICAL = 7
SP_ERR_MEMORY = 8

_d = C.c_double
_i64 = C.c_int64
_u64 = C.c_uint64
_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)
_vp = C.c_void_p


class CsrView(C.Structure):
    _fields_ = [("nrows", _i64), ("ncols", _i64), ("nnz", _i64), ("row_ptr", _pi64), ("col_idx", _pi64),
                ("values", _pd)]


class ScaledView(C.Structure):
    _fields_ = [("a", CsrView), ("at", CsrView), ("c", _pd), ("lc", _pd), ("uc", _pd), ("lv", _pd), ("uv", _pd),
                ("row_scale", _pd), ("col_scale", _pd), ("bound_scale", _d), ("obj_scale", _d)]


class LpView(C.Structure):
    _fields_ = [("a", CsrView), ("c", _pd), ("lc", _pd), ("uc", _pd), ("lv", _pd), ("uv", _pd), ("c0", _d),
                ("maximize", C.c_int)]


class PrimalStats(C.Structure):
    _fields_ = [("residual_sq", _d), ("objective", _d)]


class DualStats(C.Structure):
    _fields_ = [("residual_sq", _d), ("objective", _d)]


class RayStats(C.Structure):
    _fields_ = [("objective", _d), ("violation_sq", _d), ("norm_sq", _d)]


class Params(C.Structure):
    _fields_ = [("tolerance", _d), ("infeasibility_tolerance", _d), ("max_iterations", _u64),
                ("time_limit_s", _d), ("check_every", _u64), ("step_size_factor", _d), ("reflection", _d),
                ("restart_sufficient", _d), ("restart_necessary", _d), ("restart_artificial", _d),
                ("pid_kp", _d), ("pid_ki", _d), ("pid_kd", _d), ("pid_discount", _d),
                ("power_iterations", _i64), ("geometric_mean_passes", _i64), ("ruiz_iterations", _i64),
                ("has_pock_chambolle", C.c_int), ("pock_chambolle_alpha", _d),
                ("bound_objective_rescaling", C.c_int), ("termination", C.c_int), ("fused", C.c_int),
                ("log_every_checks", C.c_int), ("precision", C.c_int), ("spmv_autotune", C.c_int),
                ("gpu_scaling", C.c_int), ("reorder_columns", C.c_int), ("mixed_stall_checks", _i64),
                ("column_panels", C.c_int), ("hot_table", C.c_int), ("reorder_rows", C.c_int)]


class Solution(C.Structure):
    _fields_ = [("status", C.c_int), ("objective", _d), ("dual_objective", _d), ("primal_residual", _d),
                ("dual_residual", _d), ("gap", _d), ("iterations", _u64), ("restarts", _u64),
                ("solve_time_s", _d), ("total_time_s", _d), ("setup_time_s", _d), ("final_omega", _d),
                ("mixed_fp64_iterations", _u64)]


class ServerParams(C.Structure):
    _fields_ = [("budget_bytes", _i64), ("max_active", C.c_int), ("prep_threads", C.c_int), ("use_graphs", C.c_int),
                ("chunk_work", _i64), ("max_skipped", C.c_int)]


class JobTimes(C.Structure):
    _fields_ = [("submitted", _d), ("admitted", _d), ("completed", _d)]


class SpmvTuning(C.Structure):
    _fields_ = [("block_nnz", C.c_int), ("long_row", C.c_int), ("max_rows", C.c_int)]


class KktReport(C.Structure):
    """sp_kkt_report (check/src/lib.rs:10-25)."""
    _fields_ = [("primal_objective", _d), ("dual_objective", _d), ("constraint_violation", _d),
                ("bound_violation", _d), ("dual_violation", _d), ("relative_primal", _d), ("relative_dual", _d),
                ("relative_gap", _d)]

    def max_relative(self) -> float:
        return max(self.relative_primal, self.relative_dual, self.relative_gap)


# name -> (restype, argtypes): every function include/saddlepoint_hip.h declares.
SIGNATURES = {
    "sp_backend_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "sp_backend_destroy": (None, [_vp]),
    "sp_backend_name": (C.c_char_p, [_vp]),
    "sp_last_error": (C.c_char_p, [_vp]),
    "sp_problem_load": (C.c_int, [_vp, C.POINTER(ScaledView), C.POINTER(_vp)]),
    "sp_problem_free": (None, [_vp]),
    "sp_vec_upload": (C.c_int, [_vp, _pd, _i64, C.POINTER(_vp)]),
    "sp_vec_zeros": (C.c_int, [_vp, _i64, C.POINTER(_vp)]),
    "sp_vec_download": (C.c_int, [_vp, _vp, _pd, _i64]),
    "sp_vec_copy": (C.c_int, [_vp, _vp, _vp]),
    "sp_vec_len": (_i64, [_vp]),
    "sp_vec_free": (None, [_vp]),
    "sp_spmv": (C.c_int, [_vp, _vp, _vp, _vp]),
    "sp_spmv_t": (C.c_int, [_vp, _vp, _vp, _vp]),
    "sp_primal_step": (C.c_int, [_vp, _vp, _vp, _vp, _d, _vp, _vp]),
    "sp_dual_step": (C.c_int, [_vp, _vp, _vp, _vp, _d, _vp]),
    "sp_halpern_step": (C.c_int, [_vp, _vp, _vp, _vp, _d, _d]),
    "sp_distance_sq": (C.c_int, [_vp, _vp, _vp, _pd]),
    "sp_sub": (C.c_int, [_vp, _vp, _vp, _vp]),
    "sp_norm_sq": (C.c_int, [_vp, _vp, _pd]),
    "sp_dot": (C.c_int, [_vp, _vp, _vp, _pd]),
    "sp_scale": (C.c_int, [_vp, _vp, _d]),
    "sp_unscale_primal": (C.c_int, [_vp, _vp, _vp, _d, _vp]),
    "sp_unscale_dual": (C.c_int, [_vp, _vp, _vp, _d, _vp]),
    "sp_primal_stats": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(PrimalStats)]),
    "sp_dual_stats": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(DualStats)]),
    "sp_dual_stats_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(DualStats)]),
    "sp_step_dual_residual_sq": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _d, _pd]),
    "sp_primal_ray_stats": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(RayStats)]),
    "sp_dual_ray_stats": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(RayStats)]),
    "sp_default_params": (None, [C.POINTER(Params)]),
    "sp_solve": (C.c_int, [_vp, C.POINTER(LpView), C.POINTER(Params), C.POINTER(Solution), _pd, _pd]),
    "sp_solver_create": (C.c_int, [_vp, C.POINTER(LpView), C.POINTER(Params), C.POINTER(_vp)]),
    "sp_solver_advance": (C.c_int, [_vp, _u64, C.POINTER(C.c_int)]),
    "sp_solver_result": (C.c_int, [_vp, C.POINTER(Solution), _pd, _pd]),
    "sp_solver_free": (None, [_vp]),
    "sp_backend_profile": (C.c_int, [_vp, C.c_int]),
    "sp_backend_profile_read": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, _pd, C.POINTER(_i64)]),
    "sp_backend_sync": (C.c_int, [_vp]),
    "sp_probe": (C.c_int, [_vp, C.c_int, _i64, _i64, C.c_int, _pd]),
    "sp_gather_replay_time": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _pd]),
    "sp_solver_kernel_time": (C.c_int, [_vp, C.c_int, C.c_int, _pd]),
    "sp_solver_problem": (_vp, [_vp]),
    "sp_spmv_tuning_default": (None, [C.c_void_p]),
    "sp_spmv_tuning_for": (None, [C.POINTER(_i64), _i64, C.c_void_p]),
    "sp_problem_load_tuned": (C.c_int, [_vp, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(_vp)]),
    "sp_spmv_time": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _pd]),
    "sp_problem_set_panels": (C.c_int, [_vp, _vp, C.c_int]),
    "sp_default_server_params": (None, [C.POINTER(ServerParams)]),
    "sp_server_create": (C.c_int, [_vp, C.POINTER(ServerParams), C.POINTER(_vp)]),
    "sp_server_free": (None, [_vp]),
    "sp_server_submit": (C.c_int, [_vp, C.POINTER(LpView), C.POINTER(Params), C.POINTER(_u64)]),
    "sp_server_run_chunk": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sp_server_state": (C.c_int, [_vp, _u64, C.POINTER(C.c_int)]),
    "sp_server_result": (C.c_int, [_vp, _u64, C.POINTER(Solution), _pd, _pd, C.POINTER(JobTimes)]),
    "sp_server_forget": (C.c_int, [_vp, _u64]),
    "sp_server_clock": (_d, [_vp]),
    "sp_scale_problem": (C.c_int, [C.POINTER(LpView), C.POINTER(Params), _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd,
                                   _pd, _pd]),
    "sp_check": (C.c_int, [_vp, C.POINTER(LpView), _pd, _pd, C.c_void_p]),
    "sp_transpose_gpu": (C.c_int, [_vp, _i64, _i64, _pi64, _pi64, _pd, _pi64, _pi64, _pd]),
    "sp_lp_read_mps": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "sp_lp_view_of": (C.c_int, [_vp, C.POINTER(LpView)]),
    "sp_lp_name": (C.c_char_p, [_vp]),
    "sp_lp_free": (None, [_vp]),
    "sp_scale_problem_gpu": (C.c_int, [_vp, C.POINTER(LpView), C.POINTER(Params), _pd, _pd, _pd, _pd, _pd, _pd, _pd,
                                       _pd, _pd, _pd, _pd]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists() or not CODE_OBJECT.exists():
        raise RuntimeError(
            f"saddlepoint HIP backend not built ({LIB_PATH} / {CODE_OBJECT} missing): "
            "run `make -C saddlepoint_amd/csrc` or __graft_entry__.build()")
    L = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class BackendError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
