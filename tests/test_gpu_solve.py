"""GPU solves against the oracle: for each instance both reach the same
status, the GPU answer passes the independent KKT check, and the objectives
agree to the tolerance (mirrors crates/saddlepoint-cuda/tests/netlib.rs:31-74).

Beyond the reference's test: iteration counts must agree within 1% (the
reference claims identical counts, README.md:9-11), and the fused execution
path must be bitwise identical to the one-kernel-per-method path.
"""
import numpy as np
import pytest

from conftest import netlib, reference_csv
from oracle import oracle as O
from saddlepoint_amd import generate
from saddlepoint_amd.backend import default_params, solve

pytestmark = pytest.mark.gpu

TOL = 1e-6
QUICK = ["afiro", "adlittle", "israel", "sctest", "shell", "etamacro", "standata", "standgub", "standmps", "forest6",
         "woodinfe"]
SLOW = ["25fv47", "80bau3b", "scrs8", "stair", "perold"]


def iterations_match(a: int, b: int) -> bool:
    return abs(a - b) <= 0.01 * max(a, b)


def run_pair(gpu, name, p, **kw):
    params = default_params(tolerance=TOL, **kw)
    c, _, _ = O.solve(p, O.default_params(tolerance=TOL, **{k: v for k, v in kw.items()
                                                             if k not in ("fused", "spmv_autotune")}))
    g = solve(gpu, p, params)
    print(f"{name:>12}: cpu {O.STATUS[c.status]} {c.iterations} it; gpu {g.status} {g.iterations} it "
          f"{g.solve_time:.3f}s", flush=True)
    assert O.STATUS[c.status] == g.status, name
    assert iterations_match(c.iterations, g.iterations), (name, c.iterations, g.iterations)
    if g.status == "Optimal":
        O.check(p, g.x, g.y).verify(2 * TOL)
        err = abs(c.objective - g.objective) / (1 + abs(c.objective) + abs(g.objective))
        assert err <= 2 * TOL, (name, c.objective, g.objective)
    return c, g


@pytest.mark.parametrize("name", QUICK)
def test_netlib_quick_gpu_matches_cpu(gpu, name):
    p = netlib(name)
    _, g = run_pair(gpu, name, p)
    status, obj = reference_csv()[name]
    if status == "optimal":
        assert abs(g.objective - obj) / (1 + abs(obj)) <= 1e-5, (g.objective, obj)
    else:
        assert g.status == "PrimalInfeasible"


@pytest.mark.parametrize("name", SLOW)
def test_netlib_slow_gpu_matches_cpu(gpu, name):
    run_pair(gpu, name, netlib(name))


@pytest.mark.parametrize("name", ["afiro", "israel", "shell", "forest6", "woodinfe"])
@pytest.mark.parametrize("termination", ["projected", "step_implied"])
def test_fused_path_is_bitwise_the_method_path(gpu, name, termination):
    """The fused kernels and the one-readback check (fused = 1) give exactly
    the iterates, counts and answer of the one-kernel-per-method path that
    follows the reference's call sequence (fused = 0)."""
    p = netlib(name)
    a = solve(gpu, p, default_params(fused=0, termination=termination))
    b = solve(gpu, p, default_params(fused=1, termination=termination))
    assert a.status == b.status and a.iterations == b.iterations and a.restarts == b.restarts
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert a.objective == b.objective


def test_step_implied_termination(gpu):
    for name in ["afiro", "adlittle", "israel"]:
        run_pair(gpu, name, netlib(name), termination="step_implied")


def test_known_optimum(gpu):
    p, opt, _, _ = generate.kat_lp(seed=5)
    g = solve(gpu, p, default_params(tolerance=1e-8))
    assert g.status == "Optimal"
    assert abs(g.objective - opt) / (1 + abs(opt)) < 1e-6, (g.objective, opt)


def test_config0_matches_oracle(gpu):
    run_pair(gpu, "config0", generate.config0())


def test_iteration_limit_and_restarts_are_reported(gpu):
    p = netlib("israel")
    g = solve(gpu, p, default_params(max_iterations=640))
    c, _, _ = O.solve(p, O.default_params(max_iterations=640))
    assert g.status == "IterationLimit" == O.STATUS[c.status]
    assert g.iterations == 640 and g.restarts == c.restarts


# ---- mixed precision: FP32 matrix/iterates between checks, FP64 checks -------------

@pytest.mark.parametrize("name", ["afiro", "adlittle", "israel", "sctest", "shell", "etamacro", "standata",
                                  "standgub", "standmps", "forest6", "woodinfe"])
def test_mixed_precision_netlib(gpu, name):
    """Same status as the FP64 oracle; an optimal answer passes the independent
    KKT check at 2e-6 and its objective is within 2e-6 of the FP64 one (both are
    within the 1e-6 tolerance of the optimum)."""
    p = netlib(name)
    c, _, _ = O.solve(p, O.default_params(tolerance=TOL))
    g = solve(gpu, p, default_params(tolerance=TOL, precision=1))
    rep = O.check(p, g.x, g.y) if g.status == "Optimal" else None
    print(f"{name:>10}: fp64 {c.iterations} it; mixed {g.status} {g.iterations} it "
          f"({g.mixed_fp64_iterations} in fp64 after switch-over); "
          f"kkt {rep.max_relative() if rep else float('nan'):.2e}", flush=True)
    assert g.status == O.STATUS[c.status]
    if g.status == "Optimal":
        rep.verify(2 * TOL)
        assert abs(c.objective - g.objective) / (1 + abs(c.objective) + abs(g.objective)) <= 2 * TOL


def test_mixed_precision_config0(gpu):
    p = generate.config0()
    c, _, _ = O.solve(p, O.default_params(tolerance=TOL))
    g = solve(gpu, p, default_params(tolerance=TOL, precision=1))
    assert g.status == "Optimal"
    O.check(p, g.x, g.y).verify(2 * TOL)
    assert abs(c.objective - g.objective) / (1 + abs(c.objective) + abs(g.objective)) <= 2 * TOL


def test_mixed_precision_follows_fp64(gpu):
    """With the displacement re-based on the FP64 iterate at every check, mixed
    precision takes the same number of iterations as FP64 on these LPs and ends
    within 1e-8 of its objective (the FP32 rounding stays far below the
    progress of a check interval)."""
    lps = [("config0", generate.config0())] + [(f"c3-{i}", p) for i, p in
                                               enumerate(generate.serving_stream(3, 12)) if i in (0, 2, 9, 11)]
    for name, p in lps:
        f = solve(gpu, p, default_params(tolerance=TOL))
        g = solve(gpu, p, default_params(tolerance=TOL, precision=1))
        gap = abs(f.objective - g.objective) / (1 + abs(f.objective))
        print(f"{name}: fp64 {f.iterations} it, mixed {g.iterations} it "
              f"({g.mixed_fp64_iterations} in fp64), objective gap {gap:.1e}", flush=True)
        assert g.status == f.status == "Optimal", name
        assert g.iterations == f.iterations and g.mixed_fp64_iterations == 0, name
        assert gap <= 1e-8, name


def test_mixed_precision_stall_window_switches_to_fp64(gpu):
    # sp_params.mixed_stall_checks = N: FP64 takes over after N checks without
    # halving the relative KKT error; the answer is still checked.
    p = generate.config0()
    g = solve(gpu, p, default_params(tolerance=TOL, precision=1, mixed_stall_checks=2))
    assert g.status == "Optimal" and 0 < g.mixed_fp64_iterations < g.iterations
    O.check(p, g.x, g.y).verify(2 * TOL)


def test_mixed_precision_needs_fused(gpu):
    from saddlepoint_amd._lib import BackendError

    with pytest.raises(BackendError):
        solve(gpu, netlib("afiro"), default_params(precision=1, fused=0))


def test_autotuned_plans_solve_like_the_oracle(gpu):
    """spmv_autotune = 1: each matrix gets the rule's plan (config0's A' takes
    a 1024-nonzero plan); the solve still matches the oracle."""
    run_pair(gpu, "config0", generate.config0(), spmv_autotune=1)
    run_pair(gpu, "israel", netlib("israel"), spmv_autotune=1)


@pytest.mark.parametrize("name", ["israel", "shell", "woodinfe", "config0"])
def test_column_relabeling_changes_storage_not_results(gpu, name):
    """reorder_columns (on by default) stores column-indexed vectors by
    decreasing column popularity; the answer comes back in the LP's order and
    agrees with the unrelabeled solve (reductions add in another order, so
    iteration counts agree within 1% rather than bitwise)."""
    p = generate.config0() if name == "config0" else netlib(name)
    a = solve(gpu, p, default_params(reorder_columns=0))
    b = solve(gpu, p, default_params(reorder_columns=1))
    print(f"{name}: plain order {a.status} {a.iterations} it; relabeled {b.status} {b.iterations} it", flush=True)
    assert a.status == b.status and iterations_match(a.iterations, b.iterations)
    if a.status == "Optimal":
        O.check(p, b.x, b.y).verify(2 * TOL)
        assert abs(a.objective - b.objective) / (1 + abs(a.objective)) <= 2 * TOL


@pytest.mark.parametrize("name", ["israel", "shell", "woodinfe", "config0"])
def test_row_relabeling_changes_storage_not_results(gpu, name):
    """reorder_rows = 1 relabels the rows (rows of more than 64 nonzeros first,
    longest first, then the others by their smallest column; by default only
    LPs of 2^20 nonzeros or more): y comes back in the LP's order and the
    solve agrees with the unrelabeled one (sums over rows add in another
    order, so iteration counts agree within 1%), and the fused path is still
    bitwise the method path."""
    p = generate.config0() if name == "config0" else netlib(name)
    a = solve(gpu, p, default_params(reorder_rows=0))
    b = solve(gpu, p, default_params(reorder_rows=1))
    c = solve(gpu, p, default_params(reorder_rows=1, fused=0))
    print(f"{name}: plain order {a.status} {a.iterations} it; relabeled {b.status} {b.iterations} it", flush=True)
    assert a.status == b.status and iterations_match(a.iterations, b.iterations)
    assert b.iterations == c.iterations and np.array_equal(b.x, c.x) and np.array_equal(b.y, c.y)
    if a.status == "Optimal":
        O.check(p, b.x, b.y).verify(2 * TOL)
        assert abs(a.objective - b.objective) / (1 + abs(a.objective)) <= 2 * TOL


def test_reorder_rows_rejects_other_values(gpu):
    from saddlepoint_amd._lib import SP_ERR_INVALID, BackendError

    with pytest.raises(BackendError) as e:
        solve(gpu, netlib("afiro"), default_params(reorder_rows=3))
    assert e.value.code == SP_ERR_INVALID


def edge_cases():
    import math

    from saddlepoint_amd.problem import CsrMatrix, LpProblem

    inf = math.inf
    return {
        # No constraint rows: the bounds alone (optimum -3 at x = (0, 3)).
        "no_rows": LpProblem(a=CsrMatrix.from_triplets(0, 2, [], [], []), c=[1.0, -1.0], lc=[], uc=[],
                             lv=[0.0, -1.0], uv=[2.0, 3.0]),
        # Rows without a nonzero.
        "no_nonzeros": LpProblem(a=CsrMatrix.from_triplets(1, 2, [], [], []), c=[1.0, 2.0], lc=[0.0], uc=[1.0],
                                 lv=[0.0, 0.0], uv=[1.0, 1.0]),
        # An empty row and an empty column among ordinary ones.
        "ragged": LpProblem(a=CsrMatrix.from_triplets(3, 3, [0, 0, 2], [0, 1, 0], [1.0, 1.0, 2.0]),
                            c=[-1.0, -2.0, 0.5], lc=[-inf, -inf, 0.0], uc=[4.0, 5.0, 6.0], lv=[0.0, 0.0, 0.0],
                            uv=[inf, 3.0, 1.0]),
        "one_variable": LpProblem(a=CsrMatrix.from_triplets(1, 1, [0], [0], [2.0]), c=[-1.0], lc=[-inf],
                                  uc=[3.0], lv=[0.0], uv=[inf]),
    }


@pytest.mark.parametrize("name", list(edge_cases()))
def test_edge_case_lps_match_the_oracle(gpu, name):
    p = edge_cases()[name]
    c, cx, _ = O.solve(p, O.default_params(tolerance=TOL))
    for kw in (dict(), dict(precision=1), dict(gpu_scaling=0), dict(reorder_rows=1)):
        g = solve(gpu, p, default_params(tolerance=TOL, **kw))
        print(f"{name} {kw}: oracle {O.STATUS[c.status]} {c.iterations} it {c.objective:.9f}; "
              f"gpu {g.status} {g.iterations} it {g.objective:.9f}", flush=True)
        assert g.status == O.STATUS[c.status] == "Optimal"
        assert abs(g.objective - c.objective) <= 2 * TOL * (1 + abs(c.objective))
        if not kw:
            assert g.iterations == c.iterations


def test_edge_case_lps_served_as_alone(gpu):
    from saddlepoint_amd.serving import Server

    items = list(edge_cases().items()) + [("afiro", netlib("afiro"))]
    s = Server(gpu)
    ids = [s.submit(p) for _, p in items]
    s.drain(ids)
    for (name, p), i in zip(items, ids):
        r = s.result(i).result
        g = solve(gpu, p)
        assert r.status == g.status and r.iterations == g.iterations, name
        assert np.array_equal(r.x, g.x) and np.array_equal(r.y, g.y), name
    s.close()


def test_fused_path_is_bitwise_the_method_path_with_long_rows(gpu):
    """Rows of several SpMV chunks (config 1's dense rows): the fused A-side
    kernel adds their chunk sums itself, in the finishing kernel's order, so
    the iterates are still bitwise those of the method path."""
    p = generate.config1(scale=0.1)
    assert np.diff(p.a.row_ptr).max() > 4096
    a = solve(gpu, p, default_params(fused=0, max_iterations=640))
    b = solve(gpu, p, default_params(fused=1, max_iterations=640))
    assert a.iterations == b.iterations == 640 and a.restarts == b.restarts
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_column_panels_solve_like_the_oracle(gpu):
    """Column panels (forced on a small wide LP: 1,024 columns per panel) keep
    the solve the oracle's: same status, iteration and restart counts over a
    capped run, iterates within 1e-8; the fused path with panels is bitwise
    the method path with panels (every A' product uses the one decomposition),
    and mixed precision runs on the FP32 panel matrix."""
    p = generate.config2(scale=0.02)
    cap = 1280
    c, cx, cy = O.solve(p, O.default_params(max_iterations=cap))
    a = solve(gpu, p, default_params(max_iterations=cap, fused=0, column_panels=10))
    b = solve(gpu, p, default_params(max_iterations=cap, fused=1, column_panels=10))
    mx = solve(gpu, p, default_params(max_iterations=cap, precision=1, column_panels=10))
    dx = np.max(np.abs(b.x - cx)) / (1 + np.max(np.abs(cx)))
    dy = np.max(np.abs(b.y - cy)) / (1 + np.max(np.abs(cy)))
    print(f"config2@0.02 panels: oracle {O.STATUS[c.status]} {c.iterations} it {c.restarts} restarts; gpu {b.status} "
          f"{b.iterations} it {b.restarts} restarts; x diff {dx:.2e} y diff {dy:.2e}; mixed {mx.status} "
          f"{mx.iterations} it", flush=True)
    assert a.status == b.status and a.iterations == b.iterations and a.restarts == b.restarts
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert b.status == O.STATUS[c.status] and b.iterations == c.iterations and b.restarts == c.restarts
    assert dx < 1e-8 and dy < 1e-8
    assert mx.iterations == cap and np.all(np.isfinite(mx.x))


@pytest.mark.parametrize("precision", [0, 1])
def test_hot_table_changes_no_bits(gpu, precision):
    """With rows relabeled long rows first (panels forced on a small wide LP),
    93% of A'ᵀ's nonzeros gather one of y's first 2,048 entries, so the
    A'-transpose kernel can keep them in LDS on a persistent grid
    (hot_table = 1 skips the size condition). The solve is bitwise the one
    without the table, and the profile shows the table's kernel ran."""
    p = generate.config2(scale=0.02)
    kw = dict(max_iterations=640, precision=precision, column_panels=10)
    a = solve(gpu, p, default_params(hot_table=0, **kw))
    gpu.profile(True)
    b = solve(gpu, p, default_params(hot_table=1, **kw))
    prof = gpu.profile_read()
    gpu.profile(False)
    main = "sp_spmv_t_primal_mixed" if precision else "sp_spmv_t_primal_halpern_f64"
    ran = sorted(k for k, (_, n) in prof.items() if n and k.startswith(main[:-4] if not precision else main))
    print(f"precision {precision}: {a.status} {a.iterations} it / {b.status} {b.iterations} it; A'T kernels {ran}",
          flush=True)
    assert any(k != main for k in ran) and main not in ran, ran
    assert a.status == b.status and a.iterations == b.iterations == 640 and a.restarts == b.restarts
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_skewed_rows_relabeled_for_the_hot_table(gpu):
    """An LP whose 2,048 longest rows hold 96% of the nonzeros (Zipf(1.5) row
    lengths, 30k rows, 1.2M nonzeros, no panels): its rows are relabeled
    longest first (reorder_rows, on from 2^20 nonzeros), and hot_table = 1
    (the size condition skipped) runs the A'-transpose side with the hot
    table. Capped, the solve keeps the oracle's status, iteration and restart
    counts with iterates within 1e-8 (the relabeling reorders sums over rows
    only); the fused path with the table is bitwise the method path under the
    same relabeling, and mixed precision runs."""
    p = generate.serving_lp(generate.seed_of(4, 7001), 20_000, 40_000, zipf_rows=1.5)
    cap = 1280
    c, cx, cy = O.solve(p, O.default_params(max_iterations=cap))
    a = solve(gpu, p, default_params(max_iterations=cap, fused=0, hot_table=1))
    gpu.profile(True)
    b = solve(gpu, p, default_params(max_iterations=cap, fused=1, hot_table=1))
    mx = solve(gpu, p, default_params(max_iterations=cap, precision=1, hot_table=1))
    prof = gpu.profile_read()
    gpu.profile(False)
    dx = np.max(np.abs(b.x - cx)) / (1 + np.max(np.abs(cx)))
    dy = np.max(np.abs(b.y - cy)) / (1 + np.max(np.abs(cy)))
    print(f"skewed m={p.nrows} nnz={p.a.nnz}: oracle {O.STATUS[c.status]} {c.iterations} it {c.restarts} restarts; "
          f"gpu {b.status} {b.iterations} it {b.restarts} restarts; x diff {dx:.2e} y diff {dy:.2e}; "
          f"mixed {mx.status} {mx.iterations} it", flush=True)
    for k in ("sp_spmv_t_primal_halpern_hot_f64", "sp_spmv_t_primal_mixed_hot"):
        assert prof.get(k, (0.0, 0))[1] > 0, (k, sorted(prof))
    assert a.status == b.status and a.iterations == b.iterations and a.restarts == b.restarts
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert b.status == O.STATUS[c.status] and b.iterations == c.iterations and b.restarts == c.restarts
    assert dx < 1e-8 and dy < 1e-8
    assert mx.iterations == cap and np.all(np.isfinite(mx.x))


def test_hot_table_rejects_other_values(gpu):
    from saddlepoint_amd._lib import SP_ERR_INVALID, BackendError

    with pytest.raises(BackendError) as e:
        solve(gpu, netlib("afiro"), default_params(hot_table=2))
    assert e.value.code == SP_ERR_INVALID
