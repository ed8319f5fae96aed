"""Parity at the benchmark sizes, through properties that do not depend on size.

* config 1 at full size (400k × 900k, ~8M nonzeros): the GPU solve and the
  oracle's, both capped at 12,800 iterations (200 checks, 11 restarts), end
  with the same status, iteration and restart counts, and iterates that agree
  far inside the solver's tolerance;
* config 2 at full size (4M × 10M, ~100M nonzeros, Zipf(1.5) row lengths):
  the same comparison capped at 1,280 iterations (20 checks, restarts), and the
  mixed-precision solve at the same cap beside FP64 (objective gap and KKT);
* SpMV at config 1 and at a quarter of config 2 (25M nonzeros, rows of up to
  ~25k nonzeros): adjointness ⟨A'x, y⟩ = ⟨x, A'ᵀy⟩, agreement with scipy's
  serial CSR product (each row summed left to right, as the CPU backend
  does), and bitwise repeatability.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sps

from oracle import oracle as O
from saddlepoint_amd import generate
from saddlepoint_amd.autotune import structural
from saddlepoint_amd.backend import default_params, solve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def config1():
    return generate.config1()


CAP = 12_800


def test_config1_capped_solve_matches_the_oracle(gpu, config1):
    import time

    p = config1
    threads = min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    t = time.perf_counter()
    c, cx, cy = O.solve(p, O.default_params(max_iterations=CAP, log_every_checks=50))
    t_oracle = time.perf_counter() - t
    t = time.perf_counter()
    g = solve(gpu, p, default_params(max_iterations=CAP))
    t_gpu = time.perf_counter() - t
    dx = np.max(np.abs(g.x - cx)) / (1 + np.max(np.abs(cx)))
    dy = np.max(np.abs(g.y - cy)) / (1 + np.max(np.abs(cy)))
    print(f"config1 capped at {CAP}: gpu {g.status} {g.iterations} it {g.restarts} restarts {t_gpu:.2f}s; oracle "
          f"{O.STATUS[c.status]} {c.iterations} it {c.restarts} restarts {t_oracle:.2f}s ({threads} threads); "
          f"x diff {dx:.2e}, y diff {dy:.2e}; objective {g.objective:.16e} vs {c.objective:.16e} "
          f"(gap {abs(g.objective - c.objective) / (1 + abs(c.objective)):.2e})", flush=True)
    assert g.status == O.STATUS[c.status] == "IterationLimit"
    assert g.iterations == c.iterations == CAP and g.restarts == c.restarts
    assert dx < 1e-8 and dy < 1e-8
    assert abs(g.objective - c.objective) <= 1e-9 * (1 + abs(c.objective))


CAP2 = 1280


def test_config2_capped_solve_matches_the_oracle(gpu):
    """Config 2 at full size against the oracle (crates/saddlepoint-cuda/tests/
    netlib.rs:31-50 at this size): same status, iteration and restart counts,
    iterates far inside the tolerance. Mixed precision at the same cap is
    reported beside FP64: its objective gap and relative KKT error at the
    same iterate count."""
    import time

    t = time.perf_counter()
    p = generate.config2()
    print(f"config2: m={p.nrows} n={p.ncols} nnz={p.a.nnz} generated in {time.perf_counter() - t:.1f}s", flush=True)
    threads = min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    t = time.perf_counter()
    c, cx, cy = O.solve(p, O.default_params(max_iterations=CAP2, log_every_checks=1))
    t_oracle = time.perf_counter() - t
    t = time.perf_counter()
    g = solve(gpu, p, default_params(max_iterations=CAP2))
    t_gpu = time.perf_counter() - t
    mx = solve(gpu, p, default_params(max_iterations=CAP2, precision=1))
    dx = np.max(np.abs(g.x - cx)) / (1 + np.max(np.abs(cx)))
    dy = np.max(np.abs(g.y - cy)) / (1 + np.max(np.abs(cy)))
    gap = abs(mx.objective - g.objective) / (1 + abs(g.objective))
    bn, cn = O.finite_bound_norm(p.lc, p.uc), O.norm2(p.c)

    def rel(r):  # the relative KKT error, as the termination test measures it (solver.rs:142-152)
        return max(r.primal_residual / (1 + bn), r.dual_residual / (1 + cn),
                   r.gap / (1 + abs(r.objective) + abs(r.dual_objective)))

    print(f"config2 capped at {CAP2}: gpu {g.status} {g.iterations} it {g.restarts} restarts {t_gpu:.2f}s; oracle "
          f"{O.STATUS[c.status]} {c.iterations} it {c.restarts} restarts {t_oracle:.1f}s ({threads} threads); "
          f"x diff {dx:.2e}, y diff {dy:.2e}; objective gap {abs(g.objective - c.objective) / (1 + abs(c.objective)):.2e}",
          flush=True)
    print(f"config2 mixed at {CAP2}: {mx.status} {mx.iterations} it {mx.restarts} restarts; objective gap vs FP64 "
          f"{gap:.2e}; KKT (max of primal, dual, gap) {rel(mx):.3e} vs FP64 {rel(g):.3e}", flush=True)
    assert g.status == O.STATUS[c.status] == "IterationLimit"
    assert g.iterations == c.iterations == CAP2 and g.restarts == c.restarts
    assert dx < 1e-8 and dy < 1e-8
    assert abs(g.objective - c.objective) <= 1e-9 * (1 + abs(c.objective))
    assert mx.status == "IterationLimit" and mx.iterations == CAP2
    assert np.all(np.isfinite(mx.x)) and np.all(np.isfinite(mx.y))


def spmv_properties(gpu, p, name):
    sp = structural(p)
    pg = gpu.load(sp)
    rng = np.random.default_rng(5)
    x, y = rng.normal(size=p.ncols), rng.normal(size=p.nrows)
    gx, gy = gpu.upload(x), gpu.upload(y)
    ax, aty = gpu.zeros(p.nrows), gpu.zeros(p.ncols)
    gpu.spmv(pg, gx, ax)
    gpu.spmv_t(pg, gy, aty)
    lhs, rhs = gpu.dot(ax, gy), gpu.dot(gx, aty)
    A = sps.csr_matrix((p.a.values, p.a.col_idx, p.a.row_ptr), shape=(p.nrows, p.ncols))
    ref_ax, ref_aty = A @ x, (A.T.tocsr()) @ y
    got_ax, got_aty = gpu.download(ax), gpu.download(aty)
    # Two summation orders of a row of length L differ by at most 2·γ_L·Σ|a x|
    # (γ_L = L·u/(1 - L·u), u = 2^-53): the rigorous bound, per row.
    u = 2.0 ** -53
    AT = A.T.tocsr()
    len_a, len_t = np.diff(A.indptr), np.diff(AT.indptr)
    bound_ax = 2.0 * len_a * u / (1 - len_a * u) * (np.abs(A) @ np.abs(x))
    bound_aty = 2.0 * len_t * u / (1 - len_t * u) * (np.abs(AT) @ np.abs(y))
    over_ax = np.max(np.abs(got_ax - ref_ax) - bound_ax)
    over_aty = np.max(np.abs(got_aty - ref_aty) - bound_aty)
    total = (np.abs(A) @ np.abs(x)).dot(np.abs(y))
    k = len_a.max() + max(p.nrows, p.ncols)
    adj = abs(lhs - rhs) / total
    print(f"{name}: nnz={p.a.nnz} max |Ax - ref| - bound {over_ax:.1e}, max |A'y - ref| - bound {over_aty:.1e}, "
          f"adjointness {adj:.1e} (bound {4 * k * u:.1e})", flush=True)
    assert over_ax <= 0 and over_aty <= 0
    assert adj <= 4 * k * u
    again = gpu.zeros(p.nrows)
    gpu.spmv(pg, gx, again)
    assert np.array_equal(gpu.download(again), got_ax)


def test_config1_spmv_properties(gpu, config1):
    spmv_properties(gpu, config1, "config1")


def test_config2_quarter_spmv_properties(gpu):
    spmv_properties(gpu, generate.config2(scale=0.25), "config2@0.25")
