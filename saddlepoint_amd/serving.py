"""Continuous-batching serving of many LPs on one GPU (``sp_server_*``).

LPs are submitted at any time; each ``run_chunk`` call runs 64 iterations of
every admitted LP (batched plain iterations, per-LP checks) and admits and
retires LPs at its edges, by HBM budget and slot count. Every LP's result is
bitwise the one it gets when served alone.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import SP_OK, BackendError, JobTimes, ServerParams, Solution, lib
from .backend import STATUS, HipBackend, SolveResult, default_params, lp_view
from .problem import LpProblem

QUEUED, PREPARED, ACTIVE, DONE, FAILED = range(5)


@dataclass
class Served:
    result: SolveResult
    submitted: float
    admitted: float
    completed: float

    @property
    def latency(self) -> float:
        return self.completed - self.submitted


class Server:
    def __init__(self, backend: HipBackend, budget_bytes: int | None = None, max_active: int | None = None,
                 prep_threads: int | None = None, use_graphs: bool | None = None, chunk_work: int | None = None,
                 max_skipped: int | None = None):
        p = ServerParams()
        lib().sp_default_server_params(C.byref(p))
        if budget_bytes is not None:
            p.budget_bytes = int(budget_bytes)
        if max_active is not None:
            p.max_active = int(max_active)
        if prep_threads is not None:
            p.prep_threads = int(prep_threads)
        if use_graphs is not None:
            p.use_graphs = int(bool(use_graphs))
        if chunk_work is not None:
            p.chunk_work = int(chunk_work)
        if max_skipped is not None:
            p.max_skipped = int(max_skipped)
        self.b = backend
        h = C.c_void_p()
        backend._ok(lib().sp_server_create(backend.h, C.byref(p), C.byref(h)))
        self.h = h
        self.shapes = {}
        backend._adopt(self)

    def close(self):
        if getattr(self, "h", None):
            lib().sp_server_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def clock(self) -> float:
        return lib().sp_server_clock(self.h)

    def submit(self, p: LpProblem, params=None, **kw) -> int:
        params = params or default_params(**kw)
        v, keep = lp_view(p)
        jid = C.c_uint64()
        self.b._ok(lib().sp_server_submit(self.h, C.byref(v), C.byref(params), C.byref(jid)))
        self.shapes[jid.value] = (p.nrows, p.ncols)
        return jid.value

    def run_chunk(self):
        """Runs one chunk; returns (active, waiting, finished so far)."""
        a, w, f = C.c_int(), C.c_int(), C.c_int()
        self.b._ok(lib().sp_server_run_chunk(self.h, C.byref(a), C.byref(w), C.byref(f)))
        return a.value, w.value, f.value

    def state(self, jid: int) -> int:
        s = C.c_int()
        self.b._ok(lib().sp_server_state(self.h, int(jid), C.byref(s)))
        return s.value

    def result(self, jid: int, with_solution: bool = True) -> Served:
        m, n = self.shapes[jid]
        sol, t = Solution(), JobTimes()
        x = np.zeros(n) if with_solution else None
        y = np.zeros(m) if with_solution else None
        px = x.ctypes.data_as(C.POINTER(C.c_double)) if with_solution else None
        py = y.ctypes.data_as(C.POINTER(C.c_double)) if with_solution else None
        rc = lib().sp_server_result(self.h, int(jid), C.byref(sol), px, py, C.byref(t))
        if rc != SP_OK:
            raise BackendError(rc, lib().sp_last_error(self.b.h).decode())
        r = SolveResult(STATUS[sol.status], x, y, sol.objective, sol.dual_objective, sol.primal_residual,
                        sol.dual_residual, sol.gap, int(sol.iterations), int(sol.restarts), sol.solve_time_s,
                        sol.total_time_s, sol.setup_time_s, sol.final_omega, int(sol.mixed_fp64_iterations))
        return Served(r, t.submitted, t.admitted, t.completed)

    def forget(self, jid: int) -> None:
        self.b._ok(lib().sp_server_forget(self.h, int(jid)))
        self.shapes.pop(jid, None)

    def drain(self, ids, max_chunks: int = 100000):
        """Runs chunks until every LP in `ids` has finished or failed."""
        pending = set(ids)
        for _ in range(max_chunks):
            pending = {j for j in pending if self.state(j) not in (DONE, FAILED)}
            if not pending:
                return
            self.run_chunk()
        raise RuntimeError("serving did not drain")
