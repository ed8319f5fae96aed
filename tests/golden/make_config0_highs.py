"""Writes config0_highs.json: the optimum of the config-0 LP from scipy's
bundled HiGHS (an independent solver), used to pin the oracle and the GPU
solve on that workload. Run from the repository root:

    python tests/golden/make_config0_highs.py
"""
import json
import sys
from pathlib import Path

import numpy as np
import scipy
import scipy.sparse as sps
from scipy.optimize import linprog

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from saddlepoint_amd import generate  # noqa: E402


def highs_objective(p):
    a = sps.csr_matrix((p.a.values, p.a.col_idx, p.a.row_ptr), shape=(p.nrows, p.ncols))
    eq = p.lc == p.uc
    ub_rows = ~eq & np.isfinite(p.uc)
    lb_rows = ~eq & np.isfinite(p.lc)
    a_ub = sps.vstack([a[ub_rows], -a[lb_rows]]).tocsr()
    b_ub = np.concatenate([p.uc[ub_rows], -p.lc[lb_rows]])
    bounds = [(None if not np.isfinite(lo) else lo, None if not np.isfinite(hi) else hi)
              for lo, hi in zip(p.lv, p.uv)]
    r = linprog(p.c, A_ub=a_ub, b_ub=b_ub, A_eq=a[eq], b_eq=p.lc[eq], bounds=bounds, method="highs")
    assert r.status == 0, r.message
    return float(r.fun) + p.c0


if __name__ == "__main__":
    p = generate.config0()
    out = {"generator": "saddlepoint_amd.generate.config0()", "m": p.nrows, "n": p.ncols, "nnz": p.a.nnz,
           "objective": highs_objective(p), "solver": f"scipy {scipy.__version__} linprog(method='highs')"}
    (Path(__file__).parent / "config0_highs.json").write_text(json.dumps(out, indent=1) + "\n")
    print(out)
