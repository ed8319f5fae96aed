"""The C-ABI library without a GPU: it loads, exports every function that
include/*.h declares, reports a missing device as SP_ERR_UNAVAILABLE, and its
host-side parts (validation, ScaledProblem) match the oracle bitwise."""
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import NETLIB, netlib
from oracle import oracle as O
from saddlepoint_amd import _lib, generate, mps
from saddlepoint_amd.backend import ScaledProblem, default_params
from saddlepoint_amd.problem import CsrMatrix, LpProblem

ROOT = Path(__file__).resolve().parents[1]
INF = float("inf")


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(sp_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_exports_every_declared_function():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_missing_device_is_unavailable():
    import ctypes as C

    h = C.c_void_p()
    rc = _lib.lib().sp_backend_create(0, C.byref(h))
    if rc == _lib.SP_OK:
        _lib.lib().sp_backend_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == _lib.SP_ERR_UNAVAILABLE
    assert _lib.lib().sp_last_error(None)


@pytest.mark.parametrize("name", ["afiro", "israel", "shell", "25fv47", "perold"])
def test_host_scaling_is_bitwise_the_oracle(name):
    p = netlib(name)
    g = ScaledProblem.new(p)
    o = O.ScaledProblem.of(p).arrays()
    assert np.array_equal(g.a.values, o["a"].values)
    assert np.array_equal(g.at.values, o["at"].values) and np.array_equal(g.at.col_idx, o["at"].col_idx)
    for k in ("c", "lc", "uc", "lv", "uv", "row_scale", "col_scale"):
        assert np.array_equal(getattr(g, k), o[k]), k
    assert g.bound_scale == o["bound_scale"] and g.obj_scale == o["obj_scale"]
    assert np.array_equal(g.primal_residual_factors(), o["primal_factor"])
    assert np.array_equal(g.dual_residual_factors(), o["dual_factor"])


@pytest.mark.parametrize("alpha", [1.0, 0.6])
def test_oracle_scaling_threads_change_nothing(alpha):
    """Above the parallel threshold the oracle folds rows and columns on
    several threads, each line in the serial order: the scaled problem is
    bitwise the one-thread result, and both are bitwise the host library's."""
    p = generate.config1(scale=0.05)
    assert p.a.nnz > 32768
    sc = O.default_scaling(pock_chambolle_alpha=alpha)
    O.set_threads(1)
    one = O.ScaledProblem.of(p, sc).arrays()
    O.set_threads(8)
    many = O.ScaledProblem.of(p, sc).arrays()
    O.set_threads(1)
    for k in ("c", "lc", "uc", "lv", "uv", "row_scale", "col_scale"):
        assert np.array_equal(one[k], many[k]), k
    assert np.array_equal(one["a"].values, many["a"].values)
    assert np.array_equal(one["at"].values, many["at"].values)
    g = ScaledProblem.new(p, default_params(pock_chambolle_alpha=alpha))
    assert np.array_equal(g.a.values, many["a"].values) and np.array_equal(g.col_scale, many["col_scale"])


def test_host_scaling_config0_and_options():
    p = generate.config0()
    for kw in ({}, {"bound_objective_rescaling": 0}, {"has_pock_chambolle": 0, "ruiz_iterations": 3}):
        g = ScaledProblem.new(p, default_params(**kw))
        o = O.ScaledProblem.of(p, O.default_scaling(**kw)).arrays()
        assert np.array_equal(g.a.values, o["a"].values)
        assert np.array_equal(g.c, o["c"]) and np.array_equal(g.col_scale, o["col_scale"])


def test_validation_errors():
    a = CsrMatrix.from_triplets(1, 2, [0, 0], [0, 1], [1.0, 1.0])
    good = dict(a=a, c=[1.0, 1.0], lc=[0.0], uc=[1.0], lv=[0.0, 0.0], uv=[1.0, 1.0])
    ScaledProblem.new(LpProblem(**good))
    for bad in ({"lv": [2.0, 0.0]}, {"uc": [float("nan")]}, {"c": [float("nan"), 1.0]}):
        with pytest.raises(_lib.BackendError) as e:
            ScaledProblem.new(LpProblem(**{**good, **bad}))
        assert e.value.code == _lib.SP_ERR_PROBLEM


# ---- problem model and MPS reader (io/src/lib.rs:324-382) ----------------------------

SAMPLE = """\
NAME          TESTLP
OBJSENSE
    MAX
ROWS
 N  COST
 L  LIM1
 G  LIM2
 E  MYEQN
 E  R4
COLUMNS
    X1        COST         1.0   LIM1         1.0
    X1        LIM2         1.0
    MARKER    'MARKER'     'INTORG'
    X2        COST         2.0   LIM1         1.0
    X2        MYEQN       -1.0
    MARKER    'MARKER'     'INTEND'
    X3        COST        -1.0   MYEQN        1.0
    X3        R4           1.0
RHS
    RHS       COST        -3.5
    RHS       LIM1         4.0   LIM2         1.0
    RHS       MYEQN        7.0   R4           2.0
RANGES
    RNG       LIM1         2.5   R4          -1.0
BOUNDS
 UP BND       X1           4.0
 MI BND       X2
 UP BND       X2           1.0
 FR BND       X3
ENDATA
"""


def test_mps_parses_all_sections():
    p = mps.read_mps(SAMPLE)
    assert p.name == "TESTLP" and p.maximize
    assert p.c.tolist() == [-1.0, -2.0, 1.0]
    assert p.c0 == -3.5
    assert p.lc.tolist() == [1.5, 1.0, 7.0, 1.0]
    assert p.uc.tolist() == [4.0, INF, 7.0, 2.0]
    assert p.lv.tolist() == [0.0, -INF, -INF]
    assert p.uv.tolist() == [4.0, 1.0, INF]
    assert p.a.nnz == 6
    cols, vals = p.a.row(2)
    assert cols.tolist() == [1, 2] and vals.tolist() == [-1.0, 1.0]
    p.validate()


def test_mps_rejects_unknown_rows():
    with pytest.raises(mps.MpsError) as e:
        mps.read_mps("NAME X\nROWS\n N obj\nCOLUMNS\n    x  nope  1\nENDATA\n")
    assert e.value.line == 5


def test_from_triplets_sums_duplicates_and_drops_zeros():
    a = CsrMatrix.from_triplets(2, 3, [1, 0, 1, 1, 0], [2, 1, 0, 2, 1], [1.0, 2.0, 3.0, 4.0, -2.0])
    assert a.row_ptr.tolist() == [0, 0, 2]
    assert a.col_idx.tolist() == [0, 2] and a.values.tolist() == [3.0, 5.0]


@pytest.mark.parametrize("name,m,n,nnz", [("afiro", 27, 32, 83), ("israel", 174, 142, 2269),
                                          ("25fv47", 821, 1571, 10400), ("80bau3b", 2262, 9799, 21002)])
def test_netlib_sizes(name, m, n, nnz):
    p = netlib(name)
    assert (p.nrows, p.ncols, p.a.nnz) == (m, n, nnz)


def test_autotuned_plan_rule_is_deterministic_and_in_range():
    import ctypes as C

    from saddlepoint_amd._lib import SpmvTuning, lib

    def plan(rp):
        t = SpmvTuning()
        rp = np.ascontiguousarray(rp, dtype=np.int64)
        lib().sp_spmv_tuning_for(rp.ctypes.data_as(C.POINTER(C.c_int64)), len(rp) - 1, C.byref(t))
        return t.block_nnz, t.long_row, t.max_rows

    d = SpmvTuning()
    lib().sp_spmv_tuning_default(C.byref(d))
    assert (d.block_nnz, d.long_row, d.max_rows) == (2048, 2048, 1024)
    rng = np.random.default_rng(3)
    for rows in (1, 10, 2000, 50_000):
        for mean in (2, 17, 60, 300):
            rp = np.concatenate([[0], np.cumsum(rng.integers(0, 2 * mean, rows))])
            a = plan(rp)
            assert a == plan(rp.copy())
            assert 1 <= a[0] <= 2048 and 1 <= a[1] <= a[0] and 1 <= a[2] <= 1024
    assert plan(np.zeros(5, np.int64)) == (2048, 2048, 1024)  # no nonzeros: the default


def _same_lp(a, b):
    assert (a.nrows, a.ncols, a.a.nnz) == (b.nrows, b.ncols, b.a.nnz)
    for f in ("row_ptr", "col_idx", "values"):
        assert np.array_equal(getattr(a.a, f), getattr(b.a, f)), f
    for f in ("c", "lc", "uc", "lv", "uv"):
        assert np.array_equal(np.asarray(getattr(a, f)), np.asarray(getattr(b, f))), f
    assert a.c0 == b.c0 and bool(a.maximize) == bool(b.maximize) and a.name == b.name


@pytest.mark.parametrize("path", sorted(NETLIB.glob("*.mps")), ids=lambda p: p.stem)
def test_native_mps_reader_is_the_python_reader(path):
    _same_lp(mps.read_mps_file_native(path), mps.read_mps_file(path))


def test_native_mps_reader_sections_and_errors(tmp_path):
    f = tmp_path / "sample.mps"
    f.write_text(SAMPLE)
    _same_lp(mps.read_mps_file_native(f), mps.read_mps_file(f))
    bad = tmp_path / "bad.mps"
    bad.write_text("NAME X\nROWS\n N obj\nCOLUMNS\n    x  nope  1\nENDATA\n")
    with pytest.raises(mps.MpsError) as e:
        mps.read_mps_file_native(bad)
    assert e.value.line == 5
    with pytest.raises(OSError):
        mps.read_mps_file_native(tmp_path / "missing.mps")


def test_shaped_lp_has_the_requested_shape():
    from saddlepoint_amd import generate

    p = generate.shaped_lp(5, 50, 400, 3000)
    assert (p.nrows, p.ncols) == (50, 400)
    assert 0.9 * 3000 <= p.a.nnz <= 3000
    assert np.diff(p.a.row_ptr).min() >= 1 and np.bincount(p.a.col_idx, minlength=400).min() >= 1


def test_default_server_params():
    import ctypes as C

    from saddlepoint_amd._lib import ServerParams, lib

    p = ServerParams()
    lib().sp_default_server_params(C.byref(p))
    assert (p.budget_bytes, p.max_active, p.prep_threads, p.use_graphs, p.chunk_work, p.max_skipped) == \
        (64 << 30, 256, 8, 1, 0, 0)


def test_default_solver_params_execution_options():
    """The execution options beyond SolverParams: fused path, FP64, the default
    plan, device scaling, column relabeling, no stall switch, and automatic
    column panels, hot tables and row relabeling."""
    from saddlepoint_amd.backend import default_params

    p = default_params()
    assert (p.fused, p.precision, p.spmv_autotune, p.gpu_scaling, p.reorder_columns, p.mixed_stall_checks,
            p.column_panels, p.hot_table, p.reorder_rows) == (1, 0, 0, 1, 1, 0, -1, -1, -1)
