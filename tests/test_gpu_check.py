"""The KKT checker on the device (sp_check) against the oracle's restatement of
crates/saddlepoint-check/src/lib.rs:62-118.

The reference's own cases (check/src/lib.rs:143-161): the tiny LP's optimum
passes at 1e-12, a suboptimal point and a wrong dual sign fail. Then every
field of the report on real solutions (Netlib, config 0, a serving LP) and on
arbitrary points (every sign branch of the dual terms taken). The device sums
in a fixed tree order instead of serially: objectives agree to 1e-12
relative, and the relative measures, which are at most O(1), to 1e-11.
"""
import math

import numpy as np
import pytest

from conftest import netlib
from oracle import oracle as O
from saddlepoint_amd import generate
from saddlepoint_amd.backend import check, default_params, solve
from saddlepoint_amd.problem import CsrMatrix, LpProblem

pytestmark = pytest.mark.gpu
INF = math.inf


def tiny() -> LpProblem:
    a = CsrMatrix.from_triplets(2, 2, [0, 0, 1, 1], [0, 1, 0, 1], [1.0, 2.0, 3.0, 1.0])
    return LpProblem(a=a, c=[-1.0, -1.0], lc=[-INF, -INF], uc=[4.0, 6.0], lv=[0.0, 0.0], uv=[INF, INF])


def agree(name, g, c):
    for f in ("primal_objective", "dual_objective"):
        a, b = getattr(g, f), getattr(c, f)
        assert abs(a - b) <= 1e-12 * (1 + abs(b)), (name, f, a, b)
    for f in ("relative_primal", "relative_dual", "relative_gap"):
        a, b = getattr(g, f), getattr(c, f)
        assert abs(a - b) <= 1e-11 * (1 + abs(b)), (name, f, a, b)
    for f in ("constraint_violation", "bound_violation", "dual_violation"):
        a, b = getattr(g, f), getattr(c, f)
        assert abs(a - b) <= 1e-9 * (1 + abs(b)), (name, f, a, b)


def test_reference_cases(gpu):
    r = check(gpu, tiny(), [1.6, 1.2], [-0.4, -0.2])
    assert r.max_relative() <= 1e-12 and abs(r.primal_objective + 2.8) < 1e-12
    assert check(gpu, tiny(), [1.0, 1.0], [-0.4, -0.2]).max_relative() > 1e-6
    assert check(gpu, tiny(), [1.6, 1.2], [0.4, 0.2]).relative_dual > 0.1


@pytest.mark.parametrize("name", ["afiro", "israel", "shell", "etamacro", "forest6", "config0", "serving"])
def test_solutions_check_as_on_the_host(gpu, name):
    p = {"config0": generate.config0, "serving": lambda: generate.serving_lp(11, m_lo=2000, m_hi=8000)}.get(
        name, lambda: netlib(name))()
    s = solve(gpu, p, default_params())
    g, c = check(gpu, p, s.x, s.y), O.check(p, s.x, s.y)
    print(f"{name}: {s.status} gpu max_relative {g.max_relative():.3e} host {c.max_relative():.3e}", flush=True)
    agree(name, g, c)


def test_arbitrary_points_check_as_on_the_host(gpu):
    rng = np.random.default_rng(8)
    for name in ("israel", "woodinfe", "config0"):
        p = generate.config0() if name == "config0" else netlib(name)
        for _ in range(3):
            x = rng.normal(0, 10, p.ncols)
            y = rng.normal(0, 1, p.nrows) * (rng.random(p.nrows) < 0.7)
            agree(name, check(gpu, p, x, y), O.check(p, x, y))


def test_empty_and_shape_errors(gpu):
    from saddlepoint_amd._lib import BackendError

    with pytest.raises(ValueError):
        check(gpu, tiny(), [1.0], [0.0, 0.0])
    bad = tiny()
    bad.c = np.array([np.nan, -1.0])
    with pytest.raises(BackendError):
        check(gpu, bad, [1.6, 1.2], [-0.4, -0.2])
