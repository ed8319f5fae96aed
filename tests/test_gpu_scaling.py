"""Preconditioning on the device (sp_params.gpu_scaling) against the host's
ScaledProblem::new (scaling.rs:53-103), which the CPU tests pin bitwise to the
oracle.

Extrema are exact in any order, the Pock-Chambolle sums add each line left to
right as the host does, and factors apply as v·(r_i·c_j) on A' and A'ᵀ alike,
so with the default α = 1 the device result is bitwise the host's. For α != 1
the device pow() may differ from the host's by an ulp: 1e-12 relative there.
"""
import numpy as np
import pytest

from conftest import netlib
from saddlepoint_amd import generate
from saddlepoint_amd.backend import ScaledProblem, default_params, solve

pytestmark = pytest.mark.gpu

FIELDS = ("c", "lc", "uc", "lv", "uv", "row_scale", "col_scale")


def cases():
    out = [(name, netlib(name)) for name in ("afiro", "israel", "shell", "25fv47", "perold", "woodinfe")]
    out.append(("config0", generate.config0()))
    out.append(("config1@0.05", generate.config1(scale=0.05)))  # dense rows of 1k-3k nonzeros
    out.append(("serving-3", generate.serving_lp(3, m_lo=2000, m_hi=20000)))
    return out


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    fin = np.isfinite(a) & np.isfinite(b)
    assert np.array_equal(np.isfinite(a), np.isfinite(b)) and np.array_equal(a[~fin], b[~fin])
    if not fin.any():
        return 0.0
    return float(np.max(np.abs(a[fin] - b[fin])) / (1.0 + np.max(np.abs(a[fin]))))


@pytest.mark.parametrize("name,p", cases(), ids=[c[0] for c in cases()])
def test_device_scaling_is_bitwise_the_host(gpu, name, p):
    h = ScaledProblem.new(p)
    d = ScaledProblem.new_on_device(gpu, p)
    print(f"{name}: nnz={p.a.nnz} a bitwise {same(h.a.values, d.a.values)} (rel {rel(h.a.values, d.a.values):.1e})",
          flush=True)
    assert same(h.a.values, d.a.values), name
    assert same(h.at.values, d.at.values), name
    for f in FIELDS:
        assert same(getattr(h, f), getattr(d, f)), (name, f)
    assert h.bound_scale == d.bound_scale and h.obj_scale == d.obj_scale


@pytest.mark.parametrize("kw", [dict(pock_chambolle_alpha=0.5), dict(pock_chambolle_alpha=1.7),
                                dict(geometric_mean_passes=0, ruiz_iterations=3),
                                dict(has_pock_chambolle=0, bound_objective_rescaling=0)])
def test_device_scaling_options(gpu, kw):
    for name in ("israel", "config0"):
        p = netlib(name) if name != "config0" else generate.config0()
        h = ScaledProblem.new(p, default_params(**kw))
        d = ScaledProblem.new_on_device(gpu, p, default_params(**kw))
        exact = kw.get("pock_chambolle_alpha", 1.0) == 1.0 or not kw.get("has_pock_chambolle", 1)
        for f in ("a",):
            a, b = h.a.values, d.a.values
            assert same(a, b) if exact else rel(a, b) < 1e-12, (name, kw)
        for f in FIELDS:
            a, b = getattr(h, f), getattr(d, f)
            assert same(a, b) if exact else rel(a, b) < 1e-12, (name, kw, f)


def test_device_scaled_solve_is_the_host_scaled_solve(gpu):
    for name in ("afiro", "israel", "shell"):
        p = netlib(name)
        a = solve(gpu, p, default_params(gpu_scaling=0))
        b = solve(gpu, p, default_params(gpu_scaling=1))
        assert a.status == b.status and a.iterations == b.iterations, name
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y), name
    p = generate.config0()
    a = solve(gpu, p, default_params(tolerance=1e-4, gpu_scaling=0))
    b = solve(gpu, p, default_params(tolerance=1e-4, gpu_scaling=1))
    assert a.iterations == b.iterations and np.array_equal(a.x, b.x)


def _random_csr(rng, m, n, nnz, empty_rows=0, empty_cols=0):
    rows = rng.integers(empty_rows, m, nnz)
    cols = rng.integers(empty_cols, n, nnz)
    from saddlepoint_amd.problem import CsrMatrix

    return CsrMatrix.from_triplets(m, n, rows, cols, rng.normal(size=nnz))


def transpose_cases():
    """Built lazily, so that collecting the CPU tests does not generate them."""
    from saddlepoint_amd.problem import CsrMatrix

    rng = np.random.default_rng(17)
    m = 50_000
    return {
        "afiro": lambda: netlib("afiro").a,
        "israel": lambda: netlib("israel").a,
        "25fv47": lambda: netlib("25fv47").a,
        "config1@0.05": lambda: generate.config1(scale=0.05).a,
        "config2@0.02": lambda: generate.config2(scale=0.02).a,  # Zipf(1.5) rows
        "empty rows and columns": lambda: _random_csr(rng, 5000, 7000, 30000, empty_rows=100, empty_cols=200),
        # One column in every row: a transpose row spanning many sort tiles.
        "dense column": lambda: CsrMatrix.from_triplets(m, 3, np.arange(m), np.full(m, 1), rng.normal(size=m)),
        # More than 2^24 columns: four radix passes.
        "2^25 columns": lambda: _random_csr(rng, 1000, 1 << 25, 200_000),
        "no nonzeros": lambda: CsrMatrix.from_triplets(4, 6, [], [], []),
    }


@pytest.mark.parametrize("name", list(transpose_cases()))
def test_device_transpose_is_the_host_transpose(gpu, name):
    from saddlepoint_amd.backend import transpose_on_device

    a = transpose_cases()[name]()
    h = a.transpose()
    d = transpose_on_device(gpu, a)
    assert np.array_equal(h.row_ptr, d.row_ptr), name
    assert np.array_equal(h.col_idx, d.col_idx), name
    assert np.array_equal(h.values, d.values), name
