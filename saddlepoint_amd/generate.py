"""Seeded synthetic LP generators for the benchmark workloads.

Every instance is feasible and bounded by construction: a hidden primal point
``x*`` inside the variable bounds fixes the row bounds, and a hidden dual point
``(y*, z*)`` with the sign pattern the bounds allow fixes ``c = Aᵀy* + z*``.

Workloads (sizes as used by the benchmark):

* ``config0``: 2,000 rows (60% equality, 40% ≤) × 3,000 columns, exactly 6
  distinct uniform rows per column, values ±10^U(-3, 3).
* ``config1``: 400k rows × 900k columns, ~8M nonzeros; Zipf(1.3) column counts
  plus 20 dense rows of 20k–60k nonzeros; values ±10^U(-4, 4); 30% of the
  variables binary-relaxed to [0, 1]. ``scale`` shrinks every dimension.
* ``config2``: 4M × 10M, ~100M nonzeros, Zipf(1.5) row lengths, uniform
  columns, values ±10^U(-3, 3). ``scale`` shrinks every dimension.
* ``serving_lp``: one LP of the serving stream (rows log-uniform, columns
  1.5–3× rows, 4–30 nonzeros per column).
* ``kat_lp``: a variant whose optimum is known in closed form (complementary
  x*, y*, z*), so cᵀx* is the optimal objective.

Bounds: 10% of columns free, 10% [0, ∞), the rest [0, U_j] with
U_j ~ U(1, 100) and x*_j ~ U(0, U_j). Equality rows have lc = uc = (Ax*)_i;
≤ rows have lc = -∞ and uc = (Ax*)_i + U(0, 1).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sps

from .problem import CsrMatrix, LpProblem

INF = np.inf


def seed_of(config: int, instance: int = 0) -> int:
    return config * 1_000_003 + instance


def _values(rng, size, lo, hi):
    mag = 10.0 ** rng.uniform(lo, hi, size)
    sign = np.where(rng.random(size) < 0.5, -1.0, 1.0)
    return sign * mag


def _csr_from_rows_cols(m, n, rows, cols, vals) -> CsrMatrix:
    """CSR with sorted columns; duplicate (row, col) pairs are summed, zeros dropped."""
    a = sps.csr_matrix((vals, (rows.astype(np.int64), cols.astype(np.int64))), shape=(m, n))
    a.sum_duplicates()
    a.sort_indices()
    a.eliminate_zeros()
    return CsrMatrix(m, n, a.indptr.astype(np.int64), a.indices.astype(np.int64), a.data.astype(np.float64))


def _spmat(a: CsrMatrix):
    return sps.csr_matrix((a.values, a.col_idx, a.row_ptr), shape=(a.nrows, a.ncols))


def _finish(rng, a: CsrMatrix, eq_frac: float, binary_frac: float = 0.0, name: str = "") -> LpProblem:
    m, n = a.nrows, a.ncols
    # Column bounds and the hidden primal point.
    kind = rng.random(n)
    free = kind < 0.10
    lower_only = (kind >= 0.10) & (kind < 0.20)
    boxed = ~(free | lower_only)
    ub = rng.uniform(1.0, 100.0, n)
    binary = np.zeros(n, bool)
    if binary_frac > 0:
        binary = boxed & (rng.random(n) < binary_frac / 0.8)
    lv = np.where(free, -INF, 0.0)
    uv = np.where(boxed, np.where(binary, 1.0, ub), INF)
    xs = rng.uniform(0.0, 1.0, n) * np.where(binary, 1.0, ub)
    xs = np.where(free, rng.normal(0.0, 10.0, n), xs)
    ax = _spmat(a) @ xs
    eq = np.zeros(m, bool)
    eq[: int(round(eq_frac * m))] = True
    lc = np.where(eq, ax, -INF)
    uc = np.where(eq, ax, ax + rng.uniform(0.0, 1.0, m))
    # Hidden dual point: y free on equality rows, y ≤ 0 on ≤ rows; z ≥ 0 on
    # columns with a finite lower bound, 0 on free ones.
    y = rng.normal(0.0, 1.0, m)
    y = np.where(eq, y, -np.abs(y))
    z = np.where(free, 0.0, np.abs(rng.normal(0.0, 1.0, n)))
    c = _spmat(a).T @ y + z
    return LpProblem(a=a, c=c, lc=lc, uc=uc, lv=lv, uv=uv, name=name)


def config0(seed: int = seed_of(0), m: int = 2000, n: int = 3000, per_col: int = 6) -> LpProblem:
    rng = np.random.default_rng(seed)
    rows = np.concatenate([rng.choice(m, per_col, replace=False) for _ in range(n)])
    cols = np.repeat(np.arange(n), per_col)
    vals = _values(rng, rows.size, -3.0, 3.0)
    a = _csr_from_rows_cols(m, n, rows, cols, vals)
    return _finish(rng, a, eq_frac=0.6, name=f"config0-{seed}")


def _zipf_counts(rng, count, a, lo, hi, target):
    """Zipf(a) counts clipped to [lo, hi], rescaled so that they sum to about
    ``target`` (the part above ``lo`` is scaled, so the floor stays ``lo``)."""
    k = np.clip(rng.zipf(a, count).astype(np.float64), lo, hi)
    extra = k - lo
    want = max(target - lo * count, 0)
    if extra.sum() > 0:
        extra *= want / extra.sum()
    k = np.floor(lo + extra + rng.random(count))
    return np.clip(k, lo, hi).astype(np.int64)


def config1(seed: int = seed_of(1), scale: float = 1.0) -> LpProblem:
    m, n = int(400_000 * scale), int(900_000 * scale)
    rng = np.random.default_rng(seed)
    # ~8M nonzeros in all: ~0.8M in the dense rows, the rest in the columns.
    counts = _zipf_counts(rng, n, 1.3, 1, m, int(7_200_000 * scale))
    cols = np.repeat(np.arange(n, dtype=np.int64), counts)
    rows = rng.integers(0, m, cols.size)
    # 20 dense rows with 20k-60k nonzeros each at uniform columns.
    ndense = 20
    dense_rows = rng.choice(m, ndense, replace=False)
    dl = np.minimum(rng.integers(int(20_000 * scale), int(60_000 * scale) + 1, ndense), n)
    drows = np.repeat(dense_rows, dl)
    dcols = np.concatenate([rng.choice(n, k, replace=False) for k in dl])
    rows = np.concatenate([rows, drows])
    cols = np.concatenate([cols, dcols])
    vals = _values(rng, rows.size, -4.0, 4.0)
    a = _csr_from_rows_cols(m, n, rows, cols, vals)
    return _finish(rng, a, eq_frac=0.6, binary_frac=0.3, name=f"config1-{scale:g}")


def config2(seed: int = seed_of(2), scale: float = 1.0) -> LpProblem:
    m, n = int(4_000_000 * scale), int(10_000_000 * scale)
    target = int(100_000_000 * scale)
    rng = np.random.default_rng(seed)
    lens = _zipf_counts(rng, m, 1.5, 1, n, target)
    rows = np.repeat(np.arange(m, dtype=np.int64), lens)
    cols = rng.integers(0, n, rows.size)
    vals = _values(rng, rows.size, -3.0, 3.0)
    a = _csr_from_rows_cols(m, n, rows, cols, vals)
    return _finish(rng, a, eq_frac=0.6, name=f"config2-{scale:g}")


def shaped_lp(seed: int, m: int, n: int, nnz: int) -> LpProblem:
    """A feasible LP of exactly m rows and n columns with about nnz nonzeros at
    uniformly random positions (every column gets nnz/n of them, every row at
    least one), values ±10^U(-3, 3), 60% equality rows: a random stand-in of a
    given size for an instance that is not available."""
    rng = np.random.default_rng(seed)
    per = max(1, nnz // n)
    cols = np.repeat(np.arange(n), per)
    extra = nnz - cols.size
    if extra > 0:
        cols = np.concatenate([cols, rng.integers(0, n, extra)])
    rows = rng.integers(0, m, cols.size)
    rows[: min(m, rows.size)] = np.arange(min(m, rows.size))  # no empty row
    a = _csr_from_rows_cols(m, n, rows, cols, _values(rng, cols.size, -3.0, 3.0))
    return _finish(rng, a, eq_frac=0.6)


def serving_lp(seed: int, m_lo: int = 1000, m_hi: int = 200_000, zipf_rows: float | None = None) -> LpProblem:
    """One LP of the serving stream: rows log-uniform in [m_lo, m_hi], columns
    U(1.5, 3)× rows, per-column nonzeros uniform in 4..30 (or Zipf row lengths
    when ``zipf_rows`` is given, for the large instances)."""
    rng = np.random.default_rng(seed)
    m = int(round(10 ** rng.uniform(np.log10(m_lo), np.log10(m_hi))))
    n = int(round(m * rng.uniform(1.5, 3.0)))
    if zipf_rows is None:
        counts = rng.integers(4, 31, n)
        cols = np.repeat(np.arange(n, dtype=np.int64), counts)
        rows = rng.integers(0, m, cols.size)
    else:
        lens = _zipf_counts(rng, m, zipf_rows, 1, n, int(17 * n))
        rows = np.repeat(np.arange(m, dtype=np.int64), lens)
        cols = rng.integers(0, n, rows.size)
    vals = _values(rng, rows.size, -3.0, 3.0)
    a = _csr_from_rows_cols(m, n, rows, cols, vals)
    return _finish(rng, a, eq_frac=0.6, name=f"serving-{seed}")


def serving_stream(config: int, count: int, start: int = 0):
    """Instances start .. start+count-1 of a serving stream: config 3 draws
    every LP from ``serving_lp``; config 4 makes 10% of them (a seeded draw per
    instance) large, with 200k-1M rows and Zipf(1.5) row lengths."""
    out = []
    for i in range(start, start + count):
        seed = seed_of(config, i)
        large = config == 4 and np.random.default_rng(seed ^ 0x5EED).random() < 0.1
        if large:
            out.append(serving_lp(seed, 200_000, 1_000_000, zipf_rows=1.5))
        else:
            out.append(serving_lp(seed))
    return out


def kat_lp(seed: int, m: int = 300, n: int = 500, per_col: int = 4):
    """An LP with a closed-form optimum; returns (problem, optimal objective).

    Half of the bounded columns sit at x*_j = 0 with z*_j > 0, the rest are
    interior with z*_j = 0; half of the ≤ rows are tight with y*_i < 0, the
    rest slack with y*_i = 0. Then (x*, y*, z*) satisfies the KKT conditions
    and cᵀx* is the optimum.
    """
    rng = np.random.default_rng(seed)
    rows = np.concatenate([rng.choice(m, per_col, replace=False) for _ in range(n)])
    cols = np.repeat(np.arange(n), per_col)
    vals = _values(rng, rows.size, -1.0, 1.0)
    a = _csr_from_rows_cols(m, n, rows, cols, vals)
    A = _spmat(a)
    ub = rng.uniform(1.0, 100.0, n)
    at_zero = rng.random(n) < 0.5
    x = np.where(at_zero, 0.0, rng.uniform(0.1, 0.9, n) * ub)
    lv = np.zeros(n)
    uv = ub
    z = np.where(at_zero, np.abs(rng.normal(0, 1, n)) + 0.1, 0.0)
    ax = A @ x
    eq = np.zeros(m, bool)
    eq[: int(0.6 * m)] = True
    tight = (~eq) & (rng.random(m) < 0.5)
    lc = np.where(eq, ax, -INF)
    uc = np.where(eq, ax, np.where(tight, ax, ax + rng.uniform(0.1, 1.0, m)))
    y = np.where(eq, rng.normal(0, 1, m), np.where(tight, -(np.abs(rng.normal(0, 1, m)) + 0.1), 0.0))
    c = A.T @ y + z
    p = LpProblem(a=a, c=c, lc=lc, uc=uc, lv=lv, uv=uv, name=f"kat-{seed}")
    return p, float(c @ x), x, y


def unbounded(p: LpProblem, pairs: int = 3, seed: int = 7) -> LpProblem:
    """``p`` made dual infeasible (primal unbounded): `pairs` pairs of new
    columns (u, v), each a copy of a random existing column of A, with
    u ∈ [0, ∞), v ∈ (-∞, 0] and c_u < c_v. Along d = e_u - e_v the rows do
    not move (A d = 0), both bounds hold for every step, and cᵀd = c_u - c_v
    < 0, so a feasible ``p`` stays feasible and becomes unbounded below. The
    new columns are not empty, so the ray runs through the rows' coupling."""
    rng = np.random.default_rng(seed)
    a = p.a
    m, n = a.nrows, a.ncols
    rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(a.row_ptr))
    r_all, c_all, v_all = [rows], [a.col_idx.astype(np.int64)], [a.values]
    c, lv, uv = list(p.c), list(p.lv), list(p.uv)
    nonempty = np.unique(a.col_idx)
    for q in range(pairs):
        src = int(nonempty[rng.integers(0, nonempty.size)])
        k = np.flatnonzero(a.col_idx == src)
        for j in (n + 2 * q, n + 2 * q + 1):
            r_all.append(rows[k])
            c_all.append(np.full(k.size, j, dtype=np.int64))
            v_all.append(a.values[k])
        cu = float(rng.uniform(-1.0, 1.0))
        c += [cu, cu + float(rng.uniform(0.5, 1.5))]
        lv += [0.0, -INF]
        uv += [INF, 0.0]
    a2 = CsrMatrix.from_triplets(m, n + 2 * pairs, np.concatenate(r_all), np.concatenate(c_all),
                                 np.concatenate(v_all))
    return LpProblem(a=a2, c=c, lc=p.lc, uc=p.uc, lv=lv, uv=uv, c0=p.c0, name=f"{p.name}+unbounded")


def maximized(p: LpProblem) -> LpProblem:
    """The maximization whose optimum is minus ``p``'s: maximize (-c)ᵀx - c0.
    ``LpProblem`` keeps c and c0 negated for a max problem (as the MPS reader
    stores them), so the arrays are ``p``'s own and only the sense changes;
    the solver reports the objective with the sign flipped back."""
    return LpProblem(a=p.a, c=p.c, lc=p.lc, uc=p.uc, lv=p.lv, uv=p.uv, c0=p.c0, maximize=True,
                     name=f"{p.name}-max")
