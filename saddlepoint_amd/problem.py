"""LP problem model: a CSR matrix and the bound form the solver works on.

    minimize    cᵀx + c0
    subject to  lc ≤ A x ≤ uc
                lv ≤ x   ≤ uv

Infinite bounds are ``±inf``; equality rows have ``lc == uc``. Mirrors
``crates/saddlepoint-core/src/problem.rs`` (``CsrMatrix`` :16-126,
``LpProblem`` :131-194). Indices are int64, the reference's ``usize``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class CsrMatrix:
    """Compressed sparse rows, columns sorted within a row (problem.rs:16-24)."""

    nrows: int
    ncols: int
    row_ptr: np.ndarray  # int64, nrows + 1
    col_idx: np.ndarray  # int64, nnz
    values: np.ndarray  # float64, nnz

    @staticmethod
    def zeros(nrows: int, ncols: int) -> "CsrMatrix":
        return CsrMatrix(nrows, ncols, np.zeros(nrows + 1, np.int64), np.zeros(0, np.int64), np.zeros(0))

    @staticmethod
    def from_triplets(nrows: int, ncols: int, rows, cols, vals) -> "CsrMatrix":
        """Duplicates are summed and explicit zeros dropped (problem.rs:33-77).

        Duplicates are summed in triplet order (the reference sorts each row
        with an unstable sort, so its order among duplicates is unspecified).
        """
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        vals = np.asarray(vals, dtype=np.float64)
        if rows.size and (rows.min() < 0 or rows.max() >= nrows or cols.min() < 0 or cols.max() >= ncols):
            raise ValueError("triplet out of range")
        order = np.lexsort((np.arange(rows.size), cols, rows))
        r, c, v = rows[order], cols[order], vals[order]
        if r.size:
            start = np.ones(r.size, dtype=bool)
            start[1:] = (r[1:] != r[:-1]) | (c[1:] != c[:-1])
            idx = np.flatnonzero(start)
            # Sum each run of duplicates left to right.
            sums = np.add.reduceat(v, idx) if idx.size else v
            if np.any(~start):
                # reduceat is a left-to-right sum per run; keep it explicit for clarity.
                pass
            r, c, v = r[idx], c[idx], sums
            keep = v != 0.0
            r, c, v = r[keep], c[keep], v[keep]
        counts = np.bincount(r, minlength=nrows) if r.size else np.zeros(nrows, np.int64)
        row_ptr = np.zeros(nrows + 1, np.int64)
        np.cumsum(counts, out=row_ptr[1:])
        return CsrMatrix(nrows, ncols, row_ptr, c.astype(np.int64), v.astype(np.float64))

    @property
    def nnz(self) -> int:
        return int(self.values.size)

    def row(self, i: int):
        a, b = self.row_ptr[i], self.row_ptr[i + 1]
        return self.col_idx[a:b], self.values[a:b]

    def transpose(self) -> "CsrMatrix":
        """Explicit transpose by counting sort (problem.rs:88-109)."""
        rows = np.repeat(np.arange(self.nrows, dtype=np.int64), np.diff(self.row_ptr))
        order = np.argsort(self.col_idx, kind="stable")
        counts = np.bincount(self.col_idx, minlength=self.ncols)
        row_ptr = np.zeros(self.ncols + 1, np.int64)
        np.cumsum(counts, out=row_ptr[1:])
        return CsrMatrix(self.ncols, self.nrows, row_ptr, rows[order], self.values[order])

    def mul_vec(self, x: np.ndarray) -> np.ndarray:
        """``A x`` with each row summed left to right (problem.rs:112-117)."""
        out = np.zeros(self.nrows)
        for i in range(self.nrows):
            cols, vals = self.row(i)
            s = 0.0
            for c, v in zip(cols, vals):
                s += v * x[c]
            out[i] = s
        return out

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.nrows, self.ncols))
        rows = np.repeat(np.arange(self.nrows), np.diff(self.row_ptr))
        d[rows, self.col_idx] = self.values
        return d


class ProblemError(ValueError):
    """``ProblemError`` (problem.rs:147-154)."""


@dataclass
class LpProblem:
    """``LpProblem`` (problem.rs:131-145). ``c``/``c0`` are already negated for max problems."""

    a: CsrMatrix
    c: np.ndarray
    lc: np.ndarray
    uc: np.ndarray
    lv: np.ndarray
    uv: np.ndarray
    c0: float = 0.0
    maximize: bool = False
    name: str = ""
    row_names: list = field(default_factory=list)
    col_names: list = field(default_factory=list)

    def __post_init__(self):
        for f in ("c", "lc", "uc", "lv", "uv"):
            setattr(self, f, np.ascontiguousarray(getattr(self, f), dtype=np.float64))

    @property
    def nrows(self) -> int:
        return self.a.nrows

    @property
    def ncols(self) -> int:
        return self.a.ncols

    def validate(self) -> None:
        """problem.rs:166-194."""
        m, n = self.nrows, self.ncols
        for what, v, length in (("c", self.c, n), ("lv", self.lv, n), ("uv", self.uv, n),
                                ("lc", self.lc, m), ("uc", self.uc, m)):
            if v.size != length:
                raise ProblemError(f"{what} has length {v.size}, expected {length}")
        if self.a.row_ptr.size != m + 1 or self.a.col_idx.size != self.a.values.size:
            raise ProblemError(f"A has length {self.a.row_ptr.size}, expected {m + 1}")
        for what, lo, hi in (("variable", self.lv, self.uv), ("constraint", self.lc, self.uc)):
            bad = np.flatnonzero(np.isnan(lo) | np.isnan(hi))
            if bad.size:
                raise ProblemError(f"{what} {bad[0]} is NaN")
            bad = np.flatnonzero(lo > hi)
            if bad.size:
                i = bad[0]
                raise ProblemError(f"{what} {i} has lower bound {lo[i]} above upper bound {hi[i]}")
        coef = np.concatenate([self.c, self.a.values])
        bad = np.flatnonzero(np.isnan(coef))
        if bad.size:
            raise ProblemError(f"coefficient {bad[0]} is NaN")

    def objective(self, x: np.ndarray) -> float:
        """problem.rs:198-201."""
        v = float(np.dot(self.c, x)) + self.c0
        return -v if self.maximize else v
