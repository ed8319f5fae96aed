"""Pins the oracle (the C restatement of the reference's CPU path) before it is
trusted as the checker of the GPU backend.

Each test is a known-answer test the reference itself holds, or a property
the mathematics fixes, chosen so that a plausible slip (a dropped term, a
wrong sign or index, a transposed operand) fails one of them:

* the reference's unit tests: tiny LP optimum and infeasibility detection
  (crates/saddlepoint-cpu/src/lib.rs:276-326), PID controller
  (crates/saddlepoint-core/src/solver.rs:486-504), scaling
  (crates/saddlepoint-core/src/scaling.rs:224-279), checker
  (crates/saddlepoint-check/src/lib.rs:143-161);
* the Netlib regression (crates/saddlepoint-cpu/tests/netlib.rs:43-110):
  status, KKT check and HiGHS objective from data/netlib/reference.csv;
* closed forms: dense brute force for SpMV, adjointness, Halpern special
  cases, clamp idempotence, exact sums, an LP with a closed-form optimum, and
  the config-0 optimum from scipy's HiGHS.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from conftest import NETLIB, netlib, reference_csv
from oracle import oracle as O
from saddlepoint_amd import generate
from saddlepoint_amd.problem import CsrMatrix, LpProblem

INF = math.inf
GOLDEN = Path(__file__).resolve().parent / "golden"

QUICK = ["afiro", "adlittle", "israel", "sctest", "shell", "etamacro", "standata", "standgub", "standmps", "forest6",
         "woodinfe"]
SLOW = ["25fv47", "80bau3b", "scrs8", "stair", "perold"]
KNOWN_HARD = ["greenbea", "klein1"]


def tiny() -> LpProblem:
    """min -x - y s.t. x + 2y ≤ 4, 3x + y ≤ 6, x, y ≥ 0: optimum (1.6, 1.2), -2.8."""
    a = CsrMatrix.from_triplets(2, 2, [0, 0, 1, 1], [0, 1, 0, 1], [1.0, 2.0, 3.0, 1.0])
    return LpProblem(a=a, c=[-1.0, -1.0], lc=[-INF, -INF], uc=[4.0, 6.0], lv=[0.0, 0.0], uv=[INF, INF])


def sample() -> LpProblem:
    """scaling.rs:207-222."""
    a = CsrMatrix.from_triplets(2, 3, [0, 0, 1, 1], [0, 1, 1, 2], [100.0, 0.5, 3.0, -0.01])
    return LpProblem(a=a, c=[1.0, -2.0, 30.0], lc=[1.0, -INF], uc=[1.0, 4.0], lv=[0.0, -1.0, -INF],
                     uv=[INF, 2.0, 5.0])


# ---- whole-solve known answers (cpu/src/lib.rs:276-326) -----------------------

def test_solves_tiny_lp():
    s, x, _ = O.solve(tiny(), tolerance=1e-8)
    assert O.STATUS[s.status] == "Optimal"
    assert abs(s.objective + 2.8) < 1e-6
    assert abs(x[0] - 1.6) < 1e-5 and abs(x[1] - 1.2) < 1e-5


def test_detects_primal_infeasibility():
    p = tiny()
    p.a = CsrMatrix.from_triplets(1, 2, [0, 0], [0, 1], [1.0, 1.0])
    p.lc, p.uc, p.uv = np.array([5.0]), np.array([INF]), np.array([1.0, 1.0])
    s, _, _ = O.solve(p)
    assert O.STATUS[s.status] == "PrimalInfeasible"


def test_detects_dual_infeasibility():
    p = tiny()
    p.a = CsrMatrix.from_triplets(1, 2, [0, 0], [0, 1], [1.0, -1.0])
    p.lc, p.uc = np.array([-INF]), np.array([1.0])
    s, _, _ = O.solve(p)
    assert O.STATUS[s.status] == "DualInfeasible"


def test_limits_are_reported():
    s, _, _ = O.solve(netlib("israel"), max_iterations=64)
    assert O.STATUS[s.status] == "IterationLimit" and s.iterations == 64
    s, _, _ = O.solve(netlib("israel"), time_limit_s=0.0)
    assert O.STATUS[s.status] == "TimeLimit" and s.iterations == 64


# ---- PID (solver.rs:486-504) ---------------------------------------------------

def test_pid_moves_weight_toward_displacement_ratio():
    p = O.default_params()
    pid = O.PidState(1.0)
    w = pid.update(p, 1.0, 1e-2, 1.0, 1e-3, 1e-3)
    assert abs(math.log(w) - math.log(100.0)) < 0.1, w


def test_pid_closed_form():
    # e = ln(ω dx/dy) = ln(0.01); I = e; ω' = exp(-(0.99 e + 0.01 e)) = 100 exactly in exact arithmetic.
    p = O.default_params()
    w = O.PidState(1.0).update(p, 1.0, 1e-2, 1.0, 1e-3, 1e-3)
    assert abs(w - 100.0) < 1e-9


def test_pid_safeguard_restores_best_weight():
    p = O.default_params()
    pid = O.PidState(1.0)
    best = pid.update(p, 1.0, 1.0, 4.0, 1e-3, 1e-3)
    pid.update(p, best, 1.0, 9.0, 1e-3, 1e-1)
    assert pid.update(p, 7.0, 1e-20, 1.0, 1e-3, 1e-3) == best
    assert pid.integral == 0.0


# ---- scaling (scaling.rs:224-279) --------------------------------------------------

def test_scaled_matrix_is_r_a_c():
    p = sample()
    s = O.ScaledProblem.of(p).arrays()
    for i in range(p.nrows):
        cols, vals = p.a.row(i)
        scols, svals = s["a"].row(i)
        assert np.array_equal(cols, scols)
        for j, v, sv in zip(cols, vals, svals):
            expected = v * s["row_scale"][i] * s["col_scale"][j]
            assert abs(expected - sv) <= 1e-14 * abs(expected)


def test_unscaling_round_trips_feasibility_and_objective():
    p = sample()
    s = O.ScaledProblem.of(p).arrays()
    x = np.array([0.01, 0.3, -2.0])
    xs = x / s["col_scale"] * s["bound_scale"]
    back = xs * s["col_scale"] / s["bound_scale"]
    assert np.all(np.abs(x - back) < 1e-14)
    f = 1.0 / (s["bound_scale"] * s["obj_scale"])
    assert abs(p.c @ x - (s["c"] @ xs) * f) < 1e-12
    ax = p.a.mul_vec(x)
    asx = s["a"].mul_vec(xs)
    for i in range(2):
        assert abs(ax[i] - asx[i] * s["primal_factor"][i]) < 1e-12


def test_ruiz_equilibrates_rows_and_columns():
    p = sample()
    lp, keep = O.lp_struct(p)
    sc = O.default_scaling(geometric_mean_passes=0, ruiz_iterations=30, has_pock_chambolle=0,
                           bound_objective_rescaling=0)
    s = O.ScaledProblem(O.lib().or_scale(O.C.byref(lp), O.C.byref(sc))).arrays()
    d = np.abs(s["a"].to_dense())
    for v in list(d.max(axis=1)) + list(d.max(axis=0)):
        assert abs(v - 1.0) < 1e-3, v


def test_transpose_is_an_involution_and_sorted():
    p = netlib("israel")
    t = p.a.transpose()
    tt = t.transpose()
    assert np.array_equal(tt.row_ptr, p.a.row_ptr) and np.array_equal(tt.col_idx, p.a.col_idx)
    assert np.array_equal(tt.values, p.a.values)
    for i in range(t.nrows):
        c, _ = t.row(i)
        assert np.all(np.diff(c) > 0)


# ---- backend methods against closed forms ---------------------------------------------

@pytest.fixture(scope="module")
def scaled_israel():
    return O.ScaledProblem.of(netlib("israel"))


def test_spmv_matches_dense_brute_force(scaled_israel):
    s = scaled_israel
    d = s.arrays()["a"].to_dense()
    rng = np.random.default_rng(0)
    x, y = rng.normal(size=s.n), rng.normal(size=s.m)
    assert np.allclose(s.spmv(x), d @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(s.spmv_t(y), d.T @ y, rtol=1e-13, atol=1e-13)
    # Adjointness: <Ax, y> = <x, Aᵀy>.
    assert abs(s.spmv(x) @ y - x @ s.spmv_t(y)) <= 1e-12 * (1 + abs(x @ s.spmv_t(y)))


def test_spmv_integer_values_are_exact():
    a = CsrMatrix.from_triplets(3, 4, [0, 0, 1, 2, 2, 2], [0, 3, 1, 0, 1, 2], [2.0, -3.0, 5.0, 1.0, 1.0, 7.0])
    p = LpProblem(a=a, c=np.zeros(4), lc=np.zeros(3), uc=np.ones(3), lv=np.zeros(4), uv=np.ones(4))
    d = {"a": a, "c": p.c, "lc": p.lc, "uc": p.uc, "lv": p.lv, "uv": p.uv, "row_scale": np.ones(3),
         "col_scale": np.ones(4), "bound_scale": 1.0, "obj_scale": 1.0}
    s = O.ScaledProblem.from_arrays(d)
    assert s.spmv([1.0, 2.0, 3.0, 4.0]).tolist() == [2.0 - 12.0, 10.0, 1.0 + 2.0 + 21.0]
    assert s.spmv_t([1.0, 2.0, 3.0]).tolist() == [2.0 + 3.0, 10.0 + 3.0, 21.0, -3.0]


def test_halpern_special_cases():
    rng = np.random.default_rng(1)
    z, zt, an = rng.normal(size=50), rng.normal(size=50), rng.normal(size=50)
    # w = 1, ρ = 0: z ← z_t.   w = 1, ρ = 1: z ← 2 z_t - z.   w = 0: z ← anchor.
    assert np.array_equal(O.ScaledProblem.halpern_step(z, zt, an, 1.0, 0.0), zt)
    assert np.array_equal(O.ScaledProblem.halpern_step(z, zt, an, 1.0, 1.0), 2 * zt - z)
    assert np.array_equal(O.ScaledProblem.halpern_step(z, zt, an, 0.0, 1.0), an)


def test_primal_step_projects_and_reflects(scaled_israel):
    s = scaled_israel
    d = s.arrays()
    rng = np.random.default_rng(2)
    x, aty = rng.normal(size=s.n) * 10, rng.normal(size=s.n)
    xt, xb = s.primal_step(x, aty, 0.0)  # τ = 0: x_t = clamp(x)
    assert np.array_equal(xt, np.clip(x, d["lv"], d["uv"]))
    xt2, _ = s.primal_step(xt, aty, 0.0)  # clamp is idempotent
    assert np.array_equal(xt2, xt)
    assert np.array_equal(xb, 2 * xt - x)
    xt3, _ = s.primal_step(x, aty, 0.37)
    assert np.allclose(xt3, np.clip(x - 0.37 * (d["c"] - aty), d["lv"], d["uv"]), rtol=0, atol=1e-15)


def test_dual_step_is_the_prox_of_the_bounds(scaled_israel):
    s = scaled_israel
    d = s.arrays()
    rng = np.random.default_rng(3)
    y, ax = rng.normal(size=s.m), rng.normal(size=s.m)
    sig = 1.9
    yt = s.dual_step(y, ax, sig)
    ref = y - sig * (ax - np.clip(ax - y / sig, d["lc"], d["uc"]))
    assert np.allclose(yt, ref, rtol=0, atol=1e-14)
    # An equality row (lc == uc) gives y_t = y - σ(ax - b) exactly.
    eq = d["lc"] == d["uc"]
    assert np.allclose(yt[eq], (y - sig * (ax - d["lc"]))[eq], rtol=0, atol=1e-14)


def test_reductions_match_exact_sums(scaled_israel):
    s = scaled_israel
    rng = np.random.default_rng(4)
    a, b = rng.normal(size=40000), rng.normal(size=40000)  # above the parallel threshold

    def bound(t):  # the classical recursive-summation error bound, n·u·Σ|t|
        return t.size * 1.2e-16 * math.fsum(np.abs(t))

    for got, t in ((O.ScaledProblem.dot(a, b), a * b), (O.ScaledProblem.norm_sq(a), a * a),
                   (O.ScaledProblem.distance_sq(a, b), (a - b) ** 2)):
        assert abs(got - math.fsum(t)) <= bound(t)
        assert abs(got - math.fsum(t)) <= 1e-12 * math.fsum(np.abs(t))
    small = a[:1000]
    assert O.ScaledProblem.norm_sq(small) == sum(float(v) * float(v) for v in small)  # serial left fold


def test_dual_residual_vanishes_when_every_variable_is_bounded():
    # solver.rs:406-409: with finite lv and uv, no reduced-cost sign is forbidden.
    p = netlib("afiro")
    p.uv = np.minimum(p.uv, 1e3)
    p.lv = np.maximum(p.lv, -1e3)
    s = O.ScaledProblem.of(p)
    rng = np.random.default_rng(5)
    y = rng.normal(size=s.m)
    assert s.dual_stats(y, s.spmv_t(y)).residual_sq == 0.0


def test_stats_match_their_definitions(scaled_israel):
    s = scaled_israel
    d = s.arrays()
    rng = np.random.default_rng(6)
    x, y = rng.normal(size=s.n) * 3, rng.normal(size=s.m) * 3
    ax, aty = s.spmv(x), s.spmv_t(y)

    def excess(v, lo, hi):
        return np.where(v < lo, v - lo, np.where(v > hi, v - hi, 0.0))

    def bound_term(v, lo, hi):
        return np.where((v > 0) & np.isfinite(lo), v * np.where(np.isfinite(lo), lo, 0),
                        np.where((v < 0) & np.isfinite(hi), v * np.where(np.isfinite(hi), hi, 0), 0.0))

    def forbidden(v, lo, hi):
        return np.where(((v > 0) & (lo == -INF)) | ((v < 0) & (hi == INF)), v, 0.0)

    ps = s.primal_stats(x, ax)
    assert np.isclose(ps.residual_sq, math.fsum((excess(ax, d["lc"], d["uc"]) * d["primal_factor"]) ** 2),
                      rtol=1e-12)
    assert np.isclose(ps.objective, math.fsum(d["c"] * x), rtol=1e-12, atol=1e-12)
    ds = s.dual_stats(y, aty)
    r = d["c"] - aty
    assert np.isclose(ds.residual_sq, math.fsum((forbidden(r, d["lv"], d["uv"]) * d["dual_factor"]) ** 2),
                      rtol=1e-12)
    assert np.isclose(ds.objective, math.fsum(bound_term(y, d["lc"], d["uc"])) +
                      math.fsum(bound_term(r, d["lv"], d["uv"])), rtol=1e-12, atol=1e-12)
    dr = s.dual_ray_stats(y, aty)
    assert np.isclose(dr.violation_sq, math.fsum(forbidden(y, d["lc"], d["uc"]) ** 2) +
                      math.fsum(forbidden(-aty, d["lv"], d["uv"]) ** 2), rtol=1e-12)
    assert np.isclose(dr.norm_sq, math.fsum(y * y), rtol=1e-13)


# ---- checker (check/src/lib.rs:143-161) ------------------------------------------------

def test_checker_accepts_the_known_optimum():
    r = O.check(tiny(), [1.6, 1.2], [-0.4, -0.2])
    r.verify(1e-12)
    assert abs(r.primal_objective + 2.8) < 1e-12


def test_checker_rejects_a_suboptimal_point():
    with pytest.raises(AssertionError):
        O.check(tiny(), [1.0, 1.0], [-0.4, -0.2]).verify(1e-6)


def test_checker_rejects_wrong_dual_sign():
    assert O.check(tiny(), [1.6, 1.2], [0.4, 0.2]).relative_dual > 0.1


# ---- whole solves against external optima ------------------------------------------------

@pytest.mark.parametrize("name", QUICK)
@pytest.mark.parametrize("termination", ["projected", "step_implied"])
def test_netlib_quick(name, termination):
    """crates/saddlepoint-cpu/tests/netlib.rs:43-72."""
    p = netlib(name)
    s, x, y = O.solve(p, tolerance=1e-6, termination=termination)
    check_tol, obj_tol = (2e-6, 1e-5) if termination == "projected" else (1e-4, 1e-4)
    status, obj = reference_csv()[name]
    if status == "optimal":
        assert O.STATUS[s.status] == "Optimal", (name, s.iterations)
        O.check(p, x, y).verify(check_tol)
        assert abs(s.objective - obj) / (1 + abs(obj)) <= obj_tol, (s.objective, obj)
    else:
        assert O.STATUS[s.status] == "PrimalInfeasible"


@pytest.mark.slow
@pytest.mark.parametrize("name", SLOW)
def test_netlib_slow(name):
    p = netlib(name)
    s, x, y = O.solve(p, tolerance=1e-6)
    status, obj = reference_csv()[name]
    assert O.STATUS[s.status] == "Optimal"
    O.check(p, x, y).verify(2e-6)
    assert abs(s.objective - obj) / (1 + abs(obj)) <= 1e-5


def test_every_instance_is_classified():
    on_disk = sorted(f.stem for f in NETLIB.glob("*.mps"))
    assert sorted(QUICK + SLOW + KNOWN_HARD) == on_disk


def test_closed_form_optimum():
    p, opt, xs, ys = generate.kat_lp(seed=3)
    O.check(p, xs, ys).verify(1e-9)  # the generator's certificate is itself optimal
    s, x, y = O.solve(p, tolerance=1e-8)
    assert O.STATUS[s.status] == "Optimal"
    assert abs(s.objective - opt) / (1 + abs(opt)) < 1e-6


def test_config0_matches_highs():
    ref = json.loads((GOLDEN / "config0_highs.json").read_text())
    p = generate.config0()
    s, x, y = O.solve(p, tolerance=1e-6)
    assert O.STATUS[s.status] == "Optimal"
    O.check(p, x, y).verify(2e-6)
    assert abs(s.objective - ref["objective"]) / (1 + abs(ref["objective"])) < 1e-5
    assert ref["m"] == p.nrows and ref["n"] == p.ncols and ref["nnz"] == p.a.nnz


# ---- unbounded and maximization instances (solver.rs:291-309, 424-430) ------------------

def _highs(p):
    """scipy's bundled HiGHS on the LP as stated (c and c0 as stored): (status, optimum)."""
    from scipy.optimize import linprog

    A = p.a.to_dense()
    aub, bub, aeq, beq = [], [], [], []
    for i in range(p.nrows):
        if p.lc[i] == p.uc[i]:
            aeq.append(A[i]), beq.append(p.lc[i])
            continue
        if p.uc[i] < INF:
            aub.append(A[i]), bub.append(p.uc[i])
        if p.lc[i] > -INF:
            aub.append(-A[i]), bub.append(-p.lc[i])
    bounds = list(zip([None if v == -INF else v for v in p.lv], [None if v == INF else v for v in p.uv]))
    r = linprog(p.c, A_ub=np.array(aub) if aub else None, b_ub=bub or None, A_eq=np.array(aeq) if aeq else None,
                b_eq=beq or None, bounds=bounds, method="highs")
    return r.status, (r.fun + p.c0 if r.status == 0 else None)


@pytest.mark.parametrize("name", ["afiro", "israel"])
def test_unbounded_construction_is_dual_infeasible(name):
    """generate.unbounded adds column pairs along which the objective falls
    without bound: HiGHS calls it unbounded (status 3) and the oracle certifies
    DualInfeasible, while the LP it started from is optimal for both."""
    base = netlib(name)
    p = generate.unbounded(base)
    assert _highs(base)[0] == 0 and _highs(p)[0] == 3
    s, _, _ = O.solve(p, tolerance=1e-6)
    assert O.STATUS[s.status] == "DualInfeasible"


def test_maximize_reports_the_maximum():
    """A max problem (c and c0 stored negated) is solved as the min problem and
    its objective reported with the sign flipped back (solver.rs:424-430):
    the maximum is minus the minimum, HiGHS's maximum of the stated
    objective, and the objective of the returned x (problem.rs:198-201)."""
    base = netlib("afiro")
    p = generate.maximized(base)
    s, x, y = O.solve(p, tolerance=1e-8)
    smin, _, _ = O.solve(base, tolerance=1e-8)
    assert O.STATUS[s.status] == "Optimal"
    assert s.objective == -smin.objective
    hs, hopt = _highs(base)
    assert hs == 0 and abs(s.objective + hopt) <= 1e-6 * (1 + abs(hopt))
    assert abs(s.objective - p.objective(x)) <= 1e-6 * (1 + abs(s.objective))
    O.check(p, x, y).verify(1e-6)


def test_mps_sample_is_infeasible():
    """The reader's OBJSENSE MAX sample (crates/saddlepoint-io/src/lib.rs:328-375)
    has no feasible point (R4 caps X3 at 2, MYEQN then forces X2 ≤ -5, and the
    range of LIM1 needs X1 + X2 ≥ 1.5 with X1 ≤ 4): HiGHS and the oracle agree."""
    from saddlepoint_amd import mps
    from test_cpu_abi import SAMPLE

    p = mps.read_mps(SAMPLE)
    assert p.maximize and _highs(p)[0] == 2
    s, _, _ = O.solve(p, tolerance=1e-6)
    assert O.STATUS[s.status] == "PrimalInfeasible"
