"""Free-format MPS reader (restates ``crates/saddlepoint-io/src/lib.rs:1-322``).

Supports NAME, OBJSENSE (inline or on the next line), ROWS, COLUMNS (integer
markers are read and ignored: the LP relaxation is returned), RHS (an RHS on
the objective row sets the constant to its negation), RANGES and BOUNDS
(UP, LO, FX, FR, MI, PL, BV, LI, UI). Only the first RHS, RANGES and BOUNDS
set is used.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

from .problem import CsrMatrix, LpProblem

INF = math.inf


class MpsError(ValueError):
    def __init__(self, line: int, msg: str):
        super().__init__(f"line {line}: {msg}")
        self.line = line


def _sense(s: str):
    s = s.upper()
    if s in ("MAX", "MAXIMIZE", "MAXIMISE"):
        return True
    if s in ("MIN", "MINIMIZE", "MINIMISE"):
        return False
    return None


def _num(s: str):
    try:
        v = float(s)
    except ValueError:
        return None
    return None if math.isnan(v) else v


def read_mps_file(path) -> LpProblem:
    path = Path(path)
    p = read_mps(path.read_text())
    if not p.name:
        p.name = path.stem
    return p


def read_mps_file_native(path) -> LpProblem:
    """The same LP through the library's C++ reader (``sp_lp_read_mps``), for
    files at MIPLIB scale; row and column names are not returned."""
    import ctypes as C

    from ._lib import SP_OK, LpView, lib

    h = C.c_void_p()
    rc = lib().sp_lp_read_mps(str(path).encode(), C.byref(h))
    if rc != SP_OK:
        msg = lib().sp_last_error(None).decode()
        if msg.startswith("line "):
            raise MpsError(int(msg.split(":", 1)[0][5:]), msg.split(": ", 1)[1])
        raise OSError(msg)
    try:
        v = LpView()
        lib().sp_lp_view_of(h, C.byref(v))
        m, n, nnz = v.a.nrows, v.a.ncols, v.a.nnz

        def arr(ptr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

        a = CsrMatrix(m, n, arr(v.a.row_ptr, m + 1, np.int64), arr(v.a.col_idx, nnz, np.int64),
                      arr(v.a.values, nnz, np.float64))
        return LpProblem(a=a, c=arr(v.c, n, np.float64), lc=arr(v.lc, m, np.float64), uc=arr(v.uc, m, np.float64),
                         lv=arr(v.lv, n, np.float64), uv=arr(v.uv, n, np.float64), c0=float(v.c0),
                         maximize=bool(v.maximize), name=lib().sp_lp_name(h).decode())
    finally:
        lib().sp_lp_free(h)


def read_mps(text: str) -> LpProblem:  # io/src/lib.rs:59-309
    name = ""
    maximize = False
    obj_row = None
    row_index: dict = {}
    row_names: list = []
    row_kind: list = []
    col_index: dict = {}
    col_names: list = []
    c: list = []
    tr, tc, tv = [], [], []
    rhs: list = []
    rng: list = []
    c0 = 0.0
    lv: list = []
    uv: list = []
    ignored = set()
    rhs_set = ranges_set = bounds_set = None
    section = None
    for lineno, line in enumerate(text.splitlines(), start=1):
        if line.startswith("*") or not line.strip():
            continue
        f = line.split()
        if not line.startswith((" ", "\t")):
            head = f[0].upper()
            if head == "NAME":
                name = f[1] if len(f) > 1 else ""
                section = "name"
            elif head == "OBJSENSE":
                if len(f) > 1:
                    s = _sense(f[1])
                    if s is None:
                        raise MpsError(lineno, f"bad OBJSENSE {f[1]}")
                    maximize = s
                section = "objsense"
            elif head == "OBJSENCE":
                section = "objsense"
            elif head in ("ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS"):
                section = head.lower()
            elif head == "ENDATA":
                break
            elif head in ("OBJSENSE_MAX", "MAX", "MAXIMIZE", "MIN", "MINIMIZE") and section == "objsense":
                maximize = bool(_sense(head))
            elif head in ("QUADOBJ", "QMATRIX", "QSECTION", "CSECTION", "SOS", "INDICATORS"):
                raise MpsError(lineno, f"section {f[0]} is not supported")
            else:
                raise MpsError(lineno, f"unknown section {f[0]}")
            continue

        if section == "objsense":
            s = _sense(f[0])
            if s is None:
                raise MpsError(lineno, f"bad OBJSENSE {f[0]}")
            maximize = s
        elif section == "rows":
            if len(f) < 2:
                raise MpsError(lineno, "ROWS line needs a type and a name")
            t = f[0].upper()
            if t == "N":
                if obj_row is None:
                    obj_row = f[1]
                else:
                    ignored.add(f[1])
                continue
            if t not in ("E", "L", "G"):
                raise MpsError(lineno, f"unknown row type {t}")
            if f[1] in row_index:
                raise MpsError(lineno, f"duplicate row {f[1]}")
            row_index[f[1]] = len(row_names)
            row_names.append(f[1])
            row_kind.append(t)
            rhs.append(0.0)
            rng.append(None)
        elif section == "columns":
            if len(f) >= 3 and f[1].strip("'").upper() == "MARKER":
                mk = f[2].strip("'").upper()
                if mk not in ("INTORG", "INTEND"):
                    raise MpsError(lineno, f"unknown marker {mk}")
                continue
            if len(f) < 3 or len(f) % 2 == 0:
                raise MpsError(lineno, "COLUMNS line needs a column and row/value pairs")
            j = col_index.get(f[0])
            if j is None:
                j = len(col_names)
                col_index[f[0]] = j
                col_names.append(f[0])
                c.append(0.0)
                lv.append(0.0)
                uv.append(INF)
            for k in range(1, len(f), 2):
                v = _num(f[k + 1])
                if v is None:
                    raise MpsError(lineno, f"bad number {f[k + 1]}")
                if f[k] == obj_row:
                    c[j] += v
                elif f[k] in row_index:
                    tr.append(row_index[f[k]])
                    tc.append(j)
                    tv.append(v)
                elif f[k] not in ignored:
                    raise MpsError(lineno, f"unknown row {f[k]}")
        elif section in ("rhs", "ranges"):
            if len(f) % 2 == 1:
                st, rest = f[0], f[1:]
            else:
                st, rest = "", f
            if section == "rhs":
                if rhs_set is None:
                    rhs_set = st
                if rhs_set != st:
                    continue
            else:
                if ranges_set is None:
                    ranges_set = st
                if ranges_set != st:
                    continue
            for k in range(0, len(rest), 2):
                v = _num(rest[k + 1])
                if v is None:
                    raise MpsError(lineno, f"bad number {rest[k + 1]}")
                if rest[k] == obj_row:
                    if section == "rhs":
                        c0 = -v
                elif rest[k] in row_index:
                    i = row_index[rest[k]]
                    if section == "rhs":
                        rhs[i] = v
                    else:
                        rng[i] = v
                elif rest[k] not in ignored:
                    raise MpsError(lineno, f"unknown row {rest[k]}")
        elif section == "bounds":
            kind = f[0].upper()
            needs_value = kind not in ("FR", "MI", "PL", "BV")
            nf = len(f)
            if (nf, needs_value) in ((4, True), (3, False)):
                st, col, val = f[1], f[2], (f[3] if nf > 3 else None)
            elif (nf, needs_value) in ((3, True), (2, False)):
                st, col, val = "", f[1], (f[2] if nf > 2 else None)
            elif (nf, needs_value) == (4, False):
                st, col, val = f[1], f[2], None
            else:
                raise MpsError(lineno, "malformed BOUNDS line")
            if bounds_set is None:
                bounds_set = st
            if bounds_set != st:
                continue
            if col not in col_index:
                raise MpsError(lineno, f"unknown column {col}")
            j = col_index[col]
            if val is None:
                v = 0.0
            else:
                v = _num(val)
                if v is None:
                    raise MpsError(lineno, f"bad number {val}")
            if kind in ("UP", "UI"):
                uv[j] = v
                # Classic MPS rule: a negative upper bound with the default
                # lower bound makes the lower bound -inf (io/src/lib.rs:241-245).
                if v < 0.0 and lv[j] == 0.0:
                    lv[j] = -INF
            elif kind in ("LO", "LI"):
                lv[j] = v
            elif kind == "FX":
                lv[j] = v
                uv[j] = v
            elif kind == "FR":
                lv[j] = -INF
                uv[j] = INF
            elif kind == "MI":
                lv[j] = -INF
            elif kind == "PL":
                uv[j] = INF
            elif kind == "BV":
                lv[j] = 0.0
                uv[j] = 1.0
            else:
                raise MpsError(lineno, f"unsupported bound type {kind}")
        elif section == "name":
            pass
        else:
            raise MpsError(lineno, "data before the first section")

    m = len(row_names)
    lc = np.zeros(m)
    uc = np.zeros(m)
    for i in range(m):  # io/src/lib.rs:276-289
        b = rhs[i]
        r = rng[i]
        k = row_kind[i]
        if r is None:
            lo, hi = {"E": (b, b), "L": (-INF, b), "G": (b, INF)}[k]
        elif k == "E":
            lo, hi = (b, b + r) if r >= 0.0 else (b + r, b)
        elif k == "L":
            lo, hi = b - abs(r), b
        else:
            lo, hi = b, b + abs(r)
        lc[i], uc[i] = lo, hi
    n = len(col_names)
    c_arr = np.array(c, dtype=np.float64)
    if maximize:
        c_arr = -c_arr
        c0 = -c0
    a = CsrMatrix.from_triplets(m, n, tr, tc, tv)
    return LpProblem(a=a, c=c_arr, lc=lc, uc=uc, lv=np.array(lv, dtype=np.float64),
                     uv=np.array(uv, dtype=np.float64), c0=c0, maximize=maximize, name=name,
                     row_names=row_names, col_names=col_names)
