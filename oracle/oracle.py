"""ctypes front end of the oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, as the checker or as the timed CPU baseline. The oracle is
a C restatement of the reference's CPU backend, scaling, solver and checker
(see sp_oracle.h for the file:line map).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libsp_oracle.so"

_d = C.c_double
_i64 = C.c_int64
_u64 = C.c_uint64
_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)


class Csr(C.Structure):
    _fields_ = [("nrows", _i64), ("ncols", _i64), ("nnz", _i64), ("row_ptr", _pi64), ("col_idx", _pi64),
                ("values", _pd)]


class Lp(C.Structure):
    _fields_ = [("a", Csr), ("c", _pd), ("lc", _pd), ("uc", _pd), ("lv", _pd), ("uv", _pd), ("c0", _d),
                ("maximize", C.c_int)]


class ScalingParams(C.Structure):
    _fields_ = [("geometric_mean_passes", _i64), ("ruiz_iterations", _i64), ("has_pock_chambolle", C.c_int),
                ("pock_chambolle_alpha", _d), ("bound_objective_rescaling", C.c_int)]


class Scaled(C.Structure):
    _fields_ = [("a", Csr), ("at", Csr), ("c", _pd), ("lc", _pd), ("uc", _pd), ("lv", _pd), ("uv", _pd),
                ("row_scale", _pd), ("col_scale", _pd), ("bound_scale", _d), ("obj_scale", _d),
                ("primal_factor", _pd), ("dual_factor", _pd)]


class SolverParams(C.Structure):
    _fields_ = [("tolerance", _d), ("infeasibility_tolerance", _d), ("max_iterations", _u64),
                ("time_limit_s", _d), ("check_every", _u64), ("step_size_factor", _d), ("reflection", _d),
                ("restart_sufficient", _d), ("restart_necessary", _d), ("restart_artificial", _d),
                ("pid_kp", _d), ("pid_ki", _d), ("pid_kd", _d), ("pid_discount", _d),
                ("power_iterations", _i64), ("scaling", ScalingParams), ("termination", C.c_int),
                ("log_every_checks", C.c_int)]


class Solution(C.Structure):
    _fields_ = [("status", C.c_int), ("objective", _d), ("dual_objective", _d), ("primal_residual", _d),
                ("dual_residual", _d), ("gap", _d), ("iterations", _u64), ("restarts", _u64),
                ("solve_time_s", _d), ("total_time_s", _d), ("final_omega", _d), ("loop_time_s", _d),
                ("error", C.c_char * 128)]


class KktReport(C.Structure):
    _fields_ = [("primal_objective", _d), ("dual_objective", _d), ("constraint_violation", _d),
                ("bound_violation", _d), ("dual_violation", _d), ("relative_primal", _d),
                ("relative_dual", _d), ("relative_gap", _d)]

    def max_relative(self) -> float:
        return max(self.relative_primal, self.relative_dual, self.relative_gap)

    def verify(self, tol: float) -> None:
        """check/src/lib.rs:33-42."""
        if not self.max_relative() <= tol:
            raise AssertionError(f"KKT check failed at tol {tol:e}: primal {self.relative_primal:.3e}, "
                                 f"dual {self.relative_dual:.3e}, gap {self.relative_gap:.3e}")


class PrimalStats(C.Structure):
    _fields_ = [("residual_sq", _d), ("objective", _d)]


class DualStats(C.Structure):
    _fields_ = [("residual_sq", _d), ("objective", _d)]


class RayStats(C.Structure):
    _fields_ = [("objective", _d), ("violation_sq", _d), ("norm_sq", _d)]


class Pid(C.Structure):
    _fields_ = [("integral", _d), ("has_prev", C.c_int), ("prev_error", _d), ("best_weight", _d),
                ("best_balance", _d)]


STATUS = {0: "Optimal", 1: "PrimalInfeasible", 2: "DualInfeasible", 3: "IterationLimit", 4: "TimeLimit"}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    sig = {
        "or_set_threads": (None, [C.c_int]),
        "or_get_threads": (C.c_int, []),
        "or_default_scaling_params": (None, [C.POINTER(ScalingParams)]),
        "or_default_solver_params": (None, [C.POINTER(SolverParams)]),
        "or_csr_mul_vec": (None, [C.POINTER(Csr), _pd, _pd]),
        "or_scale": (C.POINTER(Scaled), [C.POINTER(Lp), C.POINTER(ScalingParams)]),
        "or_scaled_free": (None, [C.POINTER(Scaled)]),
        "or_scaled_from_arrays": (C.POINTER(Scaled), [C.POINTER(Csr), _pd, _pd, _pd, _pd, _pd, _pd, _pd, _d, _d]),
        "or_finite_bound_norm": (_d, [_pd, _pd, _i64]),
        "or_norm2": (_d, [_pd, _i64]),
        "or_spmv": (None, [C.POINTER(Scaled), _pd, _pd]),
        "or_spmv_t": (None, [C.POINTER(Scaled), _pd, _pd]),
        "or_primal_step": (None, [C.POINTER(Scaled), _pd, _pd, _d, _pd, _pd]),
        "or_dual_step": (None, [C.POINTER(Scaled), _pd, _pd, _d, _pd]),
        "or_halpern_step": (None, [_pd, _pd, _pd, _d, _d, _i64]),
        "or_distance_sq": (_d, [_pd, _pd, _i64]),
        "or_sub": (None, [_pd, _pd, _pd, _i64]),
        "or_norm_sq": (_d, [_pd, _i64]),
        "or_dot": (_d, [_pd, _pd, _i64]),
        "or_scale_vec": (None, [_pd, _d, _i64]),
        "or_primal_stats_of": (PrimalStats, [C.POINTER(Scaled), _pd, _pd]),
        "or_dual_stats_of": (DualStats, [C.POINTER(Scaled), _pd, _pd]),
        "or_step_dual_residual_sq": (_d, [C.POINTER(Scaled), _pd, _pd, _pd, _pd, _d]),
        "or_primal_ray_stats": (RayStats, [C.POINTER(Scaled), _pd, _pd]),
        "or_dual_ray_stats": (RayStats, [C.POINTER(Scaled), _pd, _pd]),
        "or_pid_init": (None, [C.POINTER(Pid), _d]),
        "or_pid_update": (_d, [C.POINTER(Pid), C.POINTER(SolverParams), _d, _d, _d, _d, _d]),
        "or_solve": (C.c_int, [C.POINTER(Lp), C.POINTER(SolverParams), C.POINTER(Solution), _pd, _pd]),
        "or_time_iterations": (_d, [C.POINTER(Lp), C.POINTER(SolverParams), _u64, _pd]),
        "or_check": (None, [C.POINTER(Lp), _pd, _pd, C.POINTER(KktReport)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(_pd)


def _pi(a: np.ndarray):
    return a.ctypes.data_as(_pi64)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64a(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


class _Keep:
    """Holds numpy arrays alive while a ctypes struct points at them."""

    def __init__(self):
        self.arrays = []

    def f(self, a):
        a = _f64(a)
        self.arrays.append(a)
        return _p(a)

    def i(self, a):
        a = _i64a(a)
        self.arrays.append(a)
        return _pi(a)


def csr_struct(keep: _Keep, m) -> Csr:
    return Csr(m.nrows, m.ncols, m.nnz, keep.i(m.row_ptr), keep.i(m.col_idx), keep.f(m.values))


def lp_struct(p):
    keep = _Keep()
    s = Lp(csr_struct(keep, p.a), keep.f(p.c), keep.f(p.lc), keep.f(p.uc), keep.f(p.lv), keep.f(p.uv),
           float(p.c0), int(bool(p.maximize)))
    return s, keep


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def default_params(**kw) -> SolverParams:
    p = SolverParams()
    lib().or_default_solver_params(C.byref(p))
    for k, v in kw.items():
        if k == "termination":
            v = {"projected": 0, "step_implied": 1}.get(v, v)
        if k in ScalingParams.__dict__ or k in [f[0] for f in ScalingParams._fields_]:
            setattr(p.scaling, k, v)
        else:
            setattr(p, k, v)
    return p


def default_scaling(**kw) -> ScalingParams:
    s = ScalingParams()
    lib().or_default_scaling_params(C.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _arr(ptr, n) -> np.ndarray:
    if n == 0:
        return np.zeros(0)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


def _arri(ptr, n) -> np.ndarray:
    if n == 0:
        return np.zeros(0, np.int64)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


class ScaledProblem:
    """An oracle-side scaled problem (ScaledProblem + CpuProblem factors)."""

    def __init__(self, handle):
        self.h = handle
        s = handle.contents
        self.m, self.n = s.a.nrows, s.a.ncols

    @staticmethod
    def of(p, scaling: ScalingParams | None = None) -> "ScaledProblem":
        lp, keep = lp_struct(p)
        sc = scaling or default_scaling()
        return ScaledProblem(lib().or_scale(C.byref(lp), C.byref(sc)))

    @staticmethod
    def from_arrays(d: dict) -> "ScaledProblem":
        keep = _Keep()
        a = d["a"]
        cs = Csr(a.nrows, a.ncols, a.nnz, keep.i(a.row_ptr), keep.i(a.col_idx), keep.f(a.values))
        h = lib().or_scaled_from_arrays(C.byref(cs), keep.f(d["c"]), keep.f(d["lc"]), keep.f(d["uc"]),
                                        keep.f(d["lv"]), keep.f(d["uv"]), keep.f(d["row_scale"]),
                                        keep.f(d["col_scale"]), float(d["bound_scale"]), float(d["obj_scale"]))
        return ScaledProblem(h)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.or_scaled_free(self.h)
            self.h = None

    def arrays(self) -> dict:
        """The scaled problem as numpy arrays (the ``load`` input)."""
        from saddlepoint_amd.problem import CsrMatrix

        s = self.h.contents
        m, n = self.m, self.n

        def csr(c):
            return CsrMatrix(c.nrows, c.ncols, _arri(c.row_ptr, c.nrows + 1), _arri(c.col_idx, c.nnz),
                             _arr(c.values, c.nnz))

        return {
            "a": csr(s.a), "at": csr(s.at), "c": _arr(s.c, n), "lc": _arr(s.lc, m), "uc": _arr(s.uc, m),
            "lv": _arr(s.lv, n), "uv": _arr(s.uv, n), "row_scale": _arr(s.row_scale, m),
            "col_scale": _arr(s.col_scale, n), "bound_scale": s.bound_scale, "obj_scale": s.obj_scale,
            "primal_factor": _arr(s.primal_factor, m), "dual_factor": _arr(s.dual_factor, n),
        }

    # ---- CpuBackend methods (cpu/src/lib.rs:137-259) ----
    def spmv(self, x):
        x = _f64(x)
        out = np.zeros(self.m)
        lib().or_spmv(self.h, _p(x), _p(out))
        return out

    def spmv_t(self, y):
        y = _f64(y)
        out = np.zeros(self.n)
        lib().or_spmv_t(self.h, _p(y), _p(out))
        return out

    def primal_step(self, x, aty, tau):
        x, aty = _f64(x), _f64(aty)
        xt, xb = np.zeros(self.n), np.zeros(self.n)
        lib().or_primal_step(self.h, _p(x), _p(aty), float(tau), _p(xt), _p(xb))
        return xt, xb

    def dual_step(self, y, ax, sigma):
        y, ax = _f64(y), _f64(ax)
        yt = np.zeros(self.m)
        lib().or_dual_step(self.h, _p(y), _p(ax), float(sigma), _p(yt))
        return yt

    @staticmethod
    def halpern_step(z, zt, anchor, w, rho):
        z = _f64(z).copy()
        zt, anchor = _f64(zt), _f64(anchor)
        lib().or_halpern_step(_p(z), _p(zt), _p(anchor), float(w), float(rho), z.size)
        return z

    @staticmethod
    def distance_sq(a, b):
        a, b = _f64(a), _f64(b)
        return lib().or_distance_sq(_p(a), _p(b), a.size)

    @staticmethod
    def sub(a, b):
        a, b = _f64(a), _f64(b)
        out = np.zeros(a.size)
        lib().or_sub(_p(a), _p(b), _p(out), a.size)
        return out

    @staticmethod
    def norm_sq(v):
        v = _f64(v)
        return lib().or_norm_sq(_p(v), v.size)

    @staticmethod
    def dot(a, b):
        a, b = _f64(a), _f64(b)
        return lib().or_dot(_p(a), _p(b), a.size)

    @staticmethod
    def scale(v, alpha):
        v = _f64(v).copy()
        lib().or_scale_vec(_p(v), float(alpha), v.size)
        return v

    def primal_stats(self, x, ax):
        x, ax = _f64(x), _f64(ax)
        return lib().or_primal_stats_of(self.h, _p(x), _p(ax))

    def dual_stats(self, y, aty):
        y, aty = _f64(y), _f64(aty)
        return lib().or_dual_stats_of(self.h, _p(y), _p(aty))

    def step_dual_residual_sq(self, x, xt, aty, aty_t, tau):
        x, xt, aty, aty_t = map(_f64, (x, xt, aty, aty_t))
        return lib().or_step_dual_residual_sq(self.h, _p(x), _p(xt), _p(aty), _p(aty_t), float(tau))

    def primal_ray_stats(self, dx, adx):
        dx, adx = _f64(dx), _f64(adx)
        return lib().or_primal_ray_stats(self.h, _p(dx), _p(adx))

    def dual_ray_stats(self, dy, atdy):
        dy, atdy = _f64(dy), _f64(atdy)
        return lib().or_dual_ray_stats(self.h, _p(dy), _p(atdy))


def solve(p, params: SolverParams | None = None, **kw):
    """``solve(CpuBackend, p, params)``; returns (Solution, x, y)."""
    params = params or default_params(**kw)
    lp, keep = lp_struct(p)
    sol = Solution()
    x = np.zeros(p.ncols)
    y = np.zeros(p.nrows)
    rc = lib().or_solve(C.byref(lp), C.byref(params), C.byref(sol), _p(x), _p(y))
    if rc != 0:
        raise RuntimeError(f"oracle solve failed ({rc}): {sol.error.decode()}")
    return sol, x, y


def time_iterations(p, iterations: int, params: SolverParams | None = None) -> tuple[float, float]:
    """Seconds spent in the loop for exactly ``iterations`` PDHG iterations
    (power iteration excluded; check work included) and the seconds spent
    scaling."""
    params = params or default_params()
    lp, keep = lp_struct(p)
    sc = C.c_double(0.0)
    t = lib().or_time_iterations(C.byref(lp), C.byref(params), int(iterations), C.byref(sc))
    return t, sc.value


def check(p, x, y) -> KktReport:
    lp, keep = lp_struct(p)
    x, y = _f64(x), _f64(y)
    r = KktReport()
    lib().or_check(C.byref(lp), _p(x), _p(y), C.byref(r))
    return r


class PidState:
    def __init__(self, omega: float):
        self.s = Pid()
        lib().or_pid_init(C.byref(self.s), float(omega))

    def update(self, params: SolverParams, omega, dx, dy, rel_primal, rel_dual) -> float:
        return lib().or_pid_update(C.byref(self.s), C.byref(params), float(omega), float(dx), float(dy),
                                   float(rel_primal), float(rel_dual))

    @property
    def integral(self) -> float:
        return self.s.integral


def finite_bound_norm(lo, hi) -> float:
    lo, hi = _f64(lo), _f64(hi)
    return lib().or_finite_bound_norm(_p(lo), _p(hi), lo.size)


def norm2(v) -> float:
    v = _f64(v)
    return lib().or_norm2(_p(v), v.size)


if os.environ.get("SP_ORACLE_THREADS"):
    set_threads(int(os.environ["SP_ORACLE_THREADS"]))
