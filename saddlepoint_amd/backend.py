"""Python face of the HIP backend: the reference's ``Backend`` trait
(``crates/saddlepoint-core/src/backend.rs:43-137``) over the C ABI, plus
``solve`` (``solver.rs:154-157``) and ``ScaledProblem`` (``scaling.rs:37-136``).

Method names, argument order and meaning follow the trait: outputs that the
trait takes as ``&mut Vector`` are passed in as ``Vector`` objects here.
Errors raise :class:`BackendError`; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from ._lib import (SP_OK, BackendError, CsrView, DualStats, KktReport, LpView, Params, PrimalStats, RayStats,
                   ScaledView, Solution, SpmvTuning, lib)
from .problem import CsrMatrix, LpProblem

STATUS = {-1: "Running", 0: "Optimal", 1: "PrimalInfeasible", 2: "DualInfeasible", 3: "IterationLimit", 4: "TimeLimit"}


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


class _Keep(list):
    def f(self, a):
        a = _f64(a)
        self.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def i(self, a):
        a = _i64(a)
        self.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int64))

    def csr(self, m: CsrMatrix) -> CsrView:
        return CsrView(m.nrows, m.ncols, m.nnz, self.i(m.row_ptr), self.i(m.col_idx), self.f(m.values))


def lp_view(p: LpProblem):
    keep = _Keep()
    v = LpView(keep.csr(p.a), keep.f(p.c), keep.f(p.lc), keep.f(p.uc), keep.f(p.lv), keep.f(p.uv), float(p.c0),
               int(bool(p.maximize)))
    return v, keep


def default_params(**kw) -> Params:
    """``SolverParams::default()`` (solver.rs:70-93) with overrides."""
    p = Params()
    lib().sp_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "termination":
            v = {"projected": 0, "step_implied": 1}.get(v, v)
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


@dataclass
class ScaledProblem:
    """``ScaledProblem`` (scaling.rs:37-50), built on the host by the library."""

    a: CsrMatrix
    at: CsrMatrix
    c: np.ndarray
    lc: np.ndarray
    uc: np.ndarray
    lv: np.ndarray
    uv: np.ndarray
    row_scale: np.ndarray
    col_scale: np.ndarray
    bound_scale: float
    obj_scale: float

    @staticmethod
    def new(p: LpProblem, params: Params | None = None) -> "ScaledProblem":
        params = params or default_params()
        v, keep = lp_view(p)
        m, n = p.nrows, p.ncols
        out = {k: np.zeros(n) for k in ("c", "lv", "uv", "col_scale")}
        out.update({k: np.zeros(m) for k in ("lc", "uc", "row_scale")})
        vals = np.zeros(p.a.nnz)
        bs, os_ = C.c_double(), C.c_double()
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        rc = lib().sp_scale_problem(C.byref(v), C.byref(params), ptr(vals), ptr(out["c"]), ptr(out["lc"]),
                                    ptr(out["uc"]), ptr(out["lv"]), ptr(out["uv"]), ptr(out["row_scale"]),
                                    ptr(out["col_scale"]), C.byref(bs), C.byref(os_))
        if rc != SP_OK:
            raise BackendError(rc, lib().sp_last_error(None).decode())
        a = CsrMatrix(m, n, p.a.row_ptr.copy(), p.a.col_idx.copy(), vals)
        return ScaledProblem(a=a, at=a.transpose(), bound_scale=bs.value, obj_scale=os_.value, **out)

    @staticmethod
    def new_on_device(backend: "HipBackend", p: LpProblem, params: Params | None = None) -> "ScaledProblem":
        """The same scaled problem, computed by the device scaling passes
        (``sp_scale_problem_gpu``)."""
        params = params or default_params()
        v, keep = lp_view(p)
        m, n = p.nrows, p.ncols
        out = {k: np.zeros(n) for k in ("c", "lv", "uv", "col_scale")}
        out.update({k: np.zeros(m) for k in ("lc", "uc", "row_scale")})
        vals, tvals = np.zeros(p.a.nnz), np.zeros(p.a.nnz)
        bs, os_ = C.c_double(), C.c_double()
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        backend._ok(lib().sp_scale_problem_gpu(backend.h, C.byref(v), C.byref(params), ptr(vals), ptr(tvals),
                                               ptr(out["c"]), ptr(out["lc"]), ptr(out["uc"]), ptr(out["lv"]),
                                               ptr(out["uv"]), ptr(out["row_scale"]), ptr(out["col_scale"]),
                                               C.byref(bs), C.byref(os_)))
        a = CsrMatrix(m, n, p.a.row_ptr.copy(), p.a.col_idx.copy(), vals)
        at = a.transpose()
        at.values = tvals
        return ScaledProblem(a=a, at=at, bound_scale=bs.value, obj_scale=os_.value, **out)

    def primal_residual_factors(self):  # scaling.rs:124-126
        return 1.0 / (self.row_scale * self.bound_scale)

    def dual_residual_factors(self):  # scaling.rs:129-131
        return 1.0 / (self.col_scale * self.obj_scale)

    def objective_factor(self) -> float:  # scaling.rs:134-136
        return 1.0 / (self.bound_scale * self.obj_scale)

    def unscale_primal(self, xs):  # scaling.rs:114-116
        return np.asarray(xs) * self.col_scale / self.bound_scale

    def unscale_dual(self, ys):  # scaling.rs:119-121
        return np.asarray(ys) * self.row_scale / self.obj_scale


class Vector:
    """A device vector owned by a backend (``Backend::Vector``)."""

    def __init__(self, backend: "HipBackend", handle):
        self._b = backend
        self.h = handle

    def __len__(self) -> int:
        return int(lib().sp_vec_len(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().sp_vec_free(self.h)
            self.h = None


class Problem:
    """A scaled problem on the device (``Backend::Problem``)."""

    def __init__(self, backend: "HipBackend", handle, m: int, n: int):
        self._b = backend
        self.h = handle
        self.m, self.n = m, n

    def free(self):
        if getattr(self, "h", None):
            lib().sp_problem_free(self.h)
            self.h = None

    def __del__(self):
        self.free()


class _BorrowedProblem:
    """A problem handle owned by someone else (a solver): never freed here."""

    def __init__(self, handle, m: int, n: int):
        self.h = handle
        self.m, self.n = m, n


class HipBackend:
    """``impl Backend`` for one MI355X device (backend.rs:43-137)."""

    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        rc = lib().sp_backend_create(int(ordinal), C.byref(h))
        if rc != SP_OK:
            raise BackendError(rc, lib().sp_last_error(None).decode())
        self.h = h
        # Servers and solvers hold the C backend's pointer (its stream, its
        # kernels): they are closed before it, also when one outlives a
        # failed caller (a traceback keeping it alive past close()).
        self._users = weakref.WeakSet()

    def _adopt(self, user) -> None:
        self._users.add(user)

    def close(self):
        if getattr(self, "h", None):
            for u in list(getattr(self, "_users", ())):
                u.close()
            lib().sp_backend_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ok(self, rc: int):
        if rc != SP_OK:
            raise BackendError(rc, lib().sp_last_error(self.h).decode())

    def name(self) -> str:
        return lib().sp_backend_name(self.h).decode()

    def load(self, s: ScaledProblem, tuning_a=None, tuning_at=None) -> Problem:
        """``Backend::load``; optionally with explicit SpMV plans for A' and
        A'ᵀ, each a (block_nnz, long_row, max_rows) tuple."""
        keep = _Keep()
        v = ScaledView(keep.csr(s.a), keep.csr(s.at), keep.f(s.c), keep.f(s.lc), keep.f(s.uc), keep.f(s.lv),
                       keep.f(s.uv), keep.f(s.row_scale), keep.f(s.col_scale), float(s.bound_scale),
                       float(s.obj_scale))
        h = C.c_void_p()
        if tuning_a is None and tuning_at is None:
            self._ok(lib().sp_problem_load(self.h, C.byref(v), C.byref(h)))
        else:
            ta = C.byref(SpmvTuning(*tuning_a)) if tuning_a is not None else None
            tt = C.byref(SpmvTuning(*tuning_at)) if tuning_at is not None else None
            self._ok(lib().sp_problem_load_tuned(self.h, C.byref(v), ta, tt, C.byref(h)))
        return Problem(self, h, s.a.nrows, s.a.ncols)

    def set_panels(self, p: Problem, shift: int) -> None:
        """Column panels of 2^shift columns for A's long rows (0: none)."""
        self._ok(lib().sp_problem_set_panels(self.h, p.h, int(shift)))

    def spmv_time(self, p: Problem, transpose: bool = False, reps: int = 20) -> float:
        """Average milliseconds of one spmv (or spmv_t) of a loaded problem."""
        ms = C.c_double()
        self._ok(lib().sp_spmv_time(self.h, p.h, int(bool(transpose)), int(reps), C.byref(ms)))
        return ms.value

    def upload(self, data) -> Vector:
        a = _f64(data)
        h = C.c_void_p()
        self._ok(lib().sp_vec_upload(self.h, a.ctypes.data_as(C.POINTER(C.c_double)), a.size, C.byref(h)))
        return Vector(self, h)

    def zeros(self, n: int) -> Vector:
        h = C.c_void_p()
        self._ok(lib().sp_vec_zeros(self.h, int(n), C.byref(h)))
        return Vector(self, h)

    def download(self, v: Vector) -> np.ndarray:
        out = np.zeros(len(v))
        self._ok(lib().sp_vec_download(self.h, v.h, out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
        return out

    def copy(self, src: Vector, dst: Vector) -> None:
        self._ok(lib().sp_vec_copy(self.h, src.h, dst.h))

    def spmv(self, p: Problem, x: Vector, out: Vector) -> None:
        self._ok(lib().sp_spmv(self.h, p.h, x.h, out.h))

    def spmv_t(self, p: Problem, y: Vector, out: Vector) -> None:
        self._ok(lib().sp_spmv_t(self.h, p.h, y.h, out.h))

    def primal_step(self, p: Problem, x: Vector, aty: Vector, tau: float, x_t: Vector, x_bar: Vector) -> None:
        self._ok(lib().sp_primal_step(self.h, p.h, x.h, aty.h, float(tau), x_t.h, x_bar.h))

    def dual_step(self, p: Problem, y: Vector, ax_bar: Vector, sigma: float, y_t: Vector) -> None:
        self._ok(lib().sp_dual_step(self.h, p.h, y.h, ax_bar.h, float(sigma), y_t.h))

    def halpern_step(self, z: Vector, z_t: Vector, anchor: Vector, w: float, rho: float) -> None:
        self._ok(lib().sp_halpern_step(self.h, z.h, z_t.h, anchor.h, float(w), float(rho)))

    def _scalar(self, fn, *args) -> float:
        out = C.c_double()
        self._ok(fn(self.h, *args, C.byref(out)))
        return out.value

    # ---- measurement (no counterpart in the reference) ----
    def profile(self, enable: bool) -> None:
        self._ok(lib().sp_backend_profile(self.h, int(bool(enable))))

    def profile_read(self) -> dict:
        """{kernel name: (total ms, launches)} since profiling was enabled."""
        out = {}
        i = 0
        name = C.create_string_buffer(128)
        while True:
            ms, n = C.c_double(), C.c_int64()
            rc = lib().sp_backend_profile_read(self.h, i, name, 128, C.byref(ms), C.byref(n))
            if rc != SP_OK:
                break
            out[name.value.decode()] = (ms.value, n.value)
            i += 1
        return out

    def sync(self) -> None:
        self._ok(lib().sp_backend_sync(self.h))

    PROBES = {"copy": 0, "triad": 1, "gather": 2, "copy_f32": 3, "triad_f32": 4, "overlap_gather": 5,
              "overlap_stream": 6, "overlap_alternate": 7, "overlap_both": 8}

    def probe(self, kind: str, n: int, table_len: int = 0, reps: int = 20) -> float:
        """Milliseconds per launch of a bandwidth probe: "copy", "triad" (f64),
        "copy_f32", "triad_f32", "gather" (random 8-byte gathers), or one of
        the "overlap_*" kinds (n gathers beside a stream of n 12-byte
        elements: gathers only, stream only, alternate workgroups, both in
        every workgroup)."""
        ms = C.c_double()
        self._ok(lib().sp_probe(self.h, self.PROBES[kind], int(n), int(table_len), int(reps), C.byref(ms)))
        return ms.value

    def gather_replay_time(self, p: Problem, transpose: bool = False, fp32: bool = False, reps: int = 20) -> float:
        """Milliseconds per launch of the gathers alone of a loaded matrix's
        SpMV (its own column indices, in the SpMV's order and tiling)."""
        ms = C.c_double()
        self._ok(lib().sp_gather_replay_time(self.h, p.h, int(bool(transpose)), int(bool(fp32)), int(reps),
                                             C.byref(ms)))
        return ms.value

    def distance_sq(self, a: Vector, b: Vector) -> float:
        return self._scalar(lib().sp_distance_sq, a.h, b.h)

    def sub(self, a: Vector, b: Vector, out: Vector) -> None:
        self._ok(lib().sp_sub(self.h, a.h, b.h, out.h))

    def norm_sq(self, v: Vector) -> float:
        return self._scalar(lib().sp_norm_sq, v.h)

    def dot(self, a: Vector, b: Vector) -> float:
        return self._scalar(lib().sp_dot, a.h, b.h)

    def scale(self, v: Vector, alpha: float) -> None:
        self._ok(lib().sp_scale(self.h, v.h, float(alpha)))

    def unscale_primal(self, xs: Vector, col_scale: Vector, bound_scale: float, out: Vector) -> None:
        """``ScaledProblem::unscale_primal`` (scaling.rs:114-116) on the device."""
        self._ok(lib().sp_unscale_primal(self.h, xs.h, col_scale.h, float(bound_scale), out.h))

    def unscale_dual(self, ys: Vector, row_scale: Vector, obj_scale: float, out: Vector) -> None:
        """``ScaledProblem::unscale_dual`` (scaling.rs:119-121) on the device."""
        self._ok(lib().sp_unscale_dual(self.h, ys.h, row_scale.h, float(obj_scale), out.h))

    def primal_stats(self, p: Problem, x: Vector, ax: Vector) -> PrimalStats:
        out = PrimalStats()
        self._ok(lib().sp_primal_stats(self.h, p.h, x.h, ax.h, C.byref(out)))
        return out

    def dual_stats(self, p: Problem, y: Vector, aty: Vector) -> DualStats:
        out = DualStats()
        self._ok(lib().sp_dual_stats(self.h, p.h, y.h, aty.h, C.byref(out)))
        return out

    def dual_stats_noalloc(self, p: Problem, y: Vector, aty: Vector) -> DualStats:
        """dual_stats through sp_dual_stats_noalloc (no device allocation)."""
        out = DualStats()
        self._ok(lib().sp_dual_stats_noalloc(self.h, p.h, y.h, aty.h, C.byref(out)))
        return out

    def step_dual_residual_sq(self, p: Problem, x: Vector, x_t: Vector, aty: Vector, aty_t: Vector,
                              tau: float) -> float:
        return self._scalar(lib().sp_step_dual_residual_sq, p.h, x.h, x_t.h, aty.h, aty_t.h, C.c_double(tau))

    def primal_ray_stats(self, p: Problem, dx: Vector, adx: Vector) -> RayStats:
        out = RayStats()
        self._ok(lib().sp_primal_ray_stats(self.h, p.h, dx.h, adx.h, C.byref(out)))
        return out

    def dual_ray_stats(self, p: Problem, dy: Vector, atdy: Vector) -> RayStats:
        out = RayStats()
        self._ok(lib().sp_dual_ray_stats(self.h, p.h, dy.h, atdy.h, C.byref(out)))
        return out


@dataclass
class SolveResult:
    """``Solution`` (solver.rs:104-121) with timings."""

    status: str
    x: np.ndarray
    y: np.ndarray
    objective: float
    dual_objective: float
    primal_residual: float
    dual_residual: float
    gap: float
    iterations: int
    restarts: int
    solve_time: float
    total_time: float
    setup_time: float
    final_omega: float
    mixed_fp64_iterations: int = 0


def transpose_on_device(backend: HipBackend, a: CsrMatrix) -> CsrMatrix:
    """``CsrMatrix::transpose`` (problem.rs:88-109) by the device's radix sort."""
    rp = np.ascontiguousarray(a.row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(a.col_idx, dtype=np.int64)
    va = np.ascontiguousarray(a.values, dtype=np.float64)
    nnz = int(rp[-1])
    trp = np.zeros(a.ncols + 1, np.int64)
    tci = np.zeros(max(nnz, 1), np.int64)
    tva = np.zeros(max(nnz, 1), np.float64)
    i64p = lambda v: v.ctypes.data_as(C.POINTER(C.c_int64))  # noqa: E731
    f64p = lambda v: v.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    backend._ok(lib().sp_transpose_gpu(backend.h, a.nrows, a.ncols, i64p(rp), i64p(ci), f64p(va), i64p(trp),
                                       i64p(tci), f64p(tva)))
    return CsrMatrix(a.ncols, a.nrows, trp, tci[:nnz], tva[:nnz])


def check(backend: HipBackend, p: LpProblem, x, y) -> KktReport:
    """``saddlepoint_check::check`` (check/src/lib.rs:62-118) on the GPU."""
    v, keep = lp_view(p)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    if x.size != p.ncols or y.size != p.nrows:
        raise ValueError("check: x and y must match the LP's columns and rows")
    r = KktReport()
    backend._ok(lib().sp_check(backend.h, C.byref(v), x.ctypes.data_as(C.POINTER(C.c_double)),
                               y.ctypes.data_as(C.POINTER(C.c_double)), C.byref(r)))
    return r


def solve(backend: HipBackend, p: LpProblem, params: Params | None = None, **kw) -> SolveResult:
    """``solve(&backend, &problem, &params)`` (solver.rs:154-157) on the GPU."""
    params = params or default_params(**kw)
    v, keep = lp_view(p)
    sol = Solution()
    x = np.zeros(p.ncols)
    y = np.zeros(p.nrows)
    rc = lib().sp_solve(backend.h, C.byref(v), C.byref(params), C.byref(sol),
                        x.ctypes.data_as(C.POINTER(C.c_double)), y.ctypes.data_as(C.POINTER(C.c_double)))
    backend._ok(rc)
    return SolveResult(STATUS[sol.status], x, y, sol.objective, sol.dual_objective, sol.primal_residual,
                       sol.dual_residual, sol.gap, int(sol.iterations), int(sol.restarts), sol.solve_time_s,
                       sol.total_time_s, sol.setup_time_s, sol.final_omega,
                       int(sol.mixed_fp64_iterations))


class Solver:
    """The driver as a steppable object (``sp_solver_*``): construction does
    solve()'s work up to the loop, including power iteration."""

    def __init__(self, backend: HipBackend, p: LpProblem, params: Params | None = None, **kw):
        self.b = backend
        self.p = p
        self.params = params or default_params(**kw)
        v, keep = lp_view(p)
        h = C.c_void_p()
        backend._ok(lib().sp_solver_create(backend.h, C.byref(v), C.byref(self.params), C.byref(h)))
        self.h = h
        backend._adopt(self)

    def advance(self, iterations: int) -> bool:
        done = C.c_int()
        self.b._ok(lib().sp_solver_advance(self.h, int(iterations), C.byref(done)))
        return bool(done.value)

    def result(self, with_solution: bool = True) -> SolveResult:
        sol = Solution()
        x = np.zeros(self.p.ncols) if with_solution else None
        y = np.zeros(self.p.nrows) if with_solution else None
        px = x.ctypes.data_as(C.POINTER(C.c_double)) if with_solution else None
        py = y.ctypes.data_as(C.POINTER(C.c_double)) if with_solution else None
        self.b._ok(lib().sp_solver_result(self.h, C.byref(sol), px, py))
        return SolveResult(STATUS[sol.status], x, y, sol.objective, sol.dual_objective, sol.primal_residual,
                           sol.dual_residual, sol.gap, int(sol.iterations), int(sol.restarts), sol.solve_time_s,
                           sol.total_time_s, sol.setup_time_s, sol.final_omega,
                       int(sol.mixed_fp64_iterations))

    def problem(self) -> "_BorrowedProblem":
        """The solver's loaded problem (owned by the solver)."""
        return _BorrowedProblem(lib().sp_solver_problem(self.h), self.p.nrows, self.p.ncols)

    def kernel_time(self, side: int, reps: int = 20) -> float:
        """Milliseconds per launch of the plain iteration's main kernel of
        side 0 (A'ᵀ) or 1 (A'), over back-to-back launches on copies."""
        ms = C.c_double()
        self.b._ok(lib().sp_solver_kernel_time(self.h, int(side), int(reps), C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().sp_solver_free(self.h)
            self.h = None

    def __del__(self):
        self.close()
