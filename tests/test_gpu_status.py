"""Dual infeasibility and maximization on the GPU, on every execution path.

The reference certifies an unbounded LP from the primal ray x_t - x_anchor
(crates/saddlepoint-core/src/solver.rs:291-309; its unit case
crates/saddlepoint-cpu/src/lib.rs:317-326) and reports a max problem's
objective with the sign flipped back (solver.rs:424-430). Here the HIP path
must reach the oracle's status and iteration count through

* fused = 0: the one-kernel-per-method sequence of the reference,
* fused = 1: the fused kernels, whose check forms the rays speculatively,
* precision = 1: mixed precision, where a near-certificate hands the LP to FP64,
* the serving loop, FP64 and mixed, where the checks run as batched phases.
"""
import math

import numpy as np
import pytest

from conftest import netlib
from oracle import oracle as O
from saddlepoint_amd import generate
from saddlepoint_amd.backend import default_params, solve
from saddlepoint_amd.problem import CsrMatrix, LpProblem
from saddlepoint_amd.serving import Server

pytestmark = pytest.mark.gpu

INF = math.inf
TOL = 1e-6


def tiny_unbounded() -> LpProblem:
    """crates/saddlepoint-cpu/src/lib.rs:317-326: min -x - y with only
    x - y ≤ 1, x, y ≥ 0: unbounded along (1, 1)."""
    a = CsrMatrix.from_triplets(1, 2, [0, 0], [0, 1], [1.0, -1.0])
    return LpProblem(a=a, c=[-1.0, -1.0], lc=[-INF], uc=[1.0], lv=[0.0, 0.0], uv=[INF, INF])


def unbounded_cases():
    return {
        "tiny": tiny_unbounded(),
        "afiro+pairs": generate.unbounded(netlib("afiro")),
        "israel+pairs": generate.unbounded(netlib("israel")),
        "config0+pairs": generate.unbounded(generate.config0()),
    }


def iterations_match(a: int, b: int) -> bool:
    return abs(a - b) <= 0.01 * max(a, b)


@pytest.mark.parametrize("name", list(unbounded_cases()))
def test_dual_infeasible_on_every_path(gpu, name):
    p = unbounded_cases()[name]
    c, _, _ = O.solve(p, O.default_params(tolerance=TOL))
    assert O.STATUS[c.status] == "DualInfeasible"
    method = solve(gpu, p, default_params(tolerance=TOL, fused=0))
    fused = solve(gpu, p, default_params(tolerance=TOL, fused=1))
    mixed = solve(gpu, p, default_params(tolerance=TOL, precision=1))
    print(f"{name}: oracle {c.iterations} it {c.restarts} restarts; method {method.status} {method.iterations}; "
          f"fused {fused.status} {fused.iterations}; mixed {mixed.status} {mixed.iterations} "
          f"({mixed.mixed_fp64_iterations} in FP64)", flush=True)
    for g in (method, fused, mixed):
        assert g.status == "DualInfeasible"
    assert iterations_match(c.iterations, method.iterations) and method.restarts == c.restarts
    # The speculative rays of the fused check are the method path's rays.
    assert fused.iterations == method.iterations and fused.restarts == method.restarts
    assert np.array_equal(fused.x, method.x) and np.array_equal(fused.y, method.y)


def test_dual_infeasible_served_as_single_solves(gpu):
    """Served with other LPs, each unbounded LP retires with the status,
    counts and bits of its single solve, in FP64 and in mixed precision."""
    cases = list(unbounded_cases().items()) + [("afiro", netlib("afiro"))]
    for precision in (0, 1):
        s = Server(gpu)
        ids = [s.submit(p, default_params(tolerance=TOL, precision=precision)) for _, p in cases]
        s.drain(ids)
        for (name, p), i in zip(cases, ids):
            r = s.result(i).result
            g = solve(gpu, p, default_params(tolerance=TOL, precision=precision))
            print(f"precision {precision} {name}: served {r.status} {r.iterations}; single {g.status} "
                  f"{g.iterations}", flush=True)
            assert r.status == g.status and r.iterations == g.iterations and r.restarts == g.restarts
            assert np.array_equal(r.x, g.x) and np.array_equal(r.y, g.y)
            if name != "afiro":
                assert r.status == "DualInfeasible"
        s.close()


@pytest.mark.parametrize("name", ["afiro", "israel", "config0"])
def test_maximize_matches_the_oracle(gpu, name):
    """max (-c)ᵀx - c0 over p's feasible set: the GPU reports the oracle's
    status, iterations and maximum, minus its own minimum of p, and the
    objective of the x it returns (problem.rs:198-201), on every path."""
    base = generate.config0() if name == "config0" else netlib(name)
    p = generate.maximized(base)
    c, _, _ = O.solve(p, O.default_params(tolerance=TOL))
    gmin = solve(gpu, base, default_params(tolerance=TOL))
    for kw in (dict(fused=0), dict(fused=1), dict(precision=1)):
        g = solve(gpu, p, default_params(tolerance=TOL, **kw))
        print(f"{name} max {kw}: oracle {O.STATUS[c.status]} {c.iterations} it {c.objective:.10g}; gpu {g.status} "
              f"{g.iterations} it {g.objective:.10g}", flush=True)
        assert g.status == O.STATUS[c.status] == "Optimal"
        assert abs(g.objective - c.objective) <= 2 * TOL * (1 + abs(c.objective))
        assert abs(g.objective - p.objective(g.x)) <= 2 * TOL * (1 + abs(g.objective))
        O.check(p, g.x, g.y).verify(2 * TOL)
        if "precision" not in kw:
            assert iterations_match(c.iterations, g.iterations)
            assert g.objective == -gmin.objective and g.dual_objective == -gmin.dual_objective


def test_maximize_served_and_mps_sample(gpu):
    """A served max LP is its single solve; the reader's OBJSENSE MAX sample
    (infeasible as written) is PrimalInfeasible on the GPU as in the oracle."""
    from saddlepoint_amd import mps
    from test_cpu_abi import SAMPLE

    sample = mps.read_mps(SAMPLE)
    c, _, _ = O.solve(sample, O.default_params(tolerance=TOL))
    items = [("afiro-max", generate.maximized(netlib("afiro"))), ("sample", sample)]
    s = Server(gpu)
    ids = [s.submit(p, default_params(tolerance=TOL)) for _, p in items]
    s.drain(ids)
    for (name, p), i in zip(items, ids):
        r = s.result(i).result
        g = solve(gpu, p, default_params(tolerance=TOL))
        print(f"{name}: served {r.status} {r.iterations} it {r.objective:.10g}; single {g.status} {g.iterations}",
              flush=True)
        assert r.status == g.status and r.iterations == g.iterations
        assert r.objective == g.objective and np.array_equal(r.x, g.x)
    assert s.result(ids[1]).result.status == O.STATUS[c.status] == "PrimalInfeasible"
    assert solve(gpu, sample, default_params(tolerance=TOL)).iterations == c.iterations
    s.close()
