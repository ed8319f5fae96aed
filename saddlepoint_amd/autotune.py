"""SpMV plan autotuner: times candidate plans on the GPU and writes the rule table.

Each matrix of a training set (both A and Aᵀ of every LP) is loaded once per
candidate plan and its SpMV timed with HIP events (``sp_spmv_time``). Matrices
are grouped into cells by size class floor(log2 nnz) and mean-row-length class
(< 8, < 32, < 128, >= 128); per cell the candidate with the least total time
wins. The cells are written as ``csrc/tuning_rule.inc``, which the library
compiles into its rule (``sp_spmv_tuning_for``), so the plan of a matrix is a
deterministic function of its row lengths, the same in every process.

Usage (on a GPU, from the repository root):
    python -m saddlepoint_amd.autotune --out <dir>
then copy ``<dir>/tuning_rule.inc`` to ``saddlepoint_amd/csrc/`` and rebuild.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import generate
from .backend import HipBackend, ScaledProblem

# block_nnz × max_rows; long_row = block_nnz (a row that does not fit one
# stream block gets workgroups of its own).
CANDIDATES = [(b, b, r) for b in (256, 512, 1024, 2048) for r in (256, 1024)]
DEFAULT = (2048, 2048, 1024)


def size_class(nnz: int) -> int:
    return max(0, int(nnz).bit_length() - 1)


def mean_class(mean: float) -> int:
    return 0 if mean < 8 else 1 if mean < 32 else 2 if mean < 128 else 3


def training_set(quick: bool = False):
    """(name, LpProblem) pairs spanning the benchmark shapes."""
    out = [("config0", generate.config0())]
    for i in range(4 if quick else 14):
        # Rows log-uniform over the serving range: a spread of sizes.
        out.append((f"config3-{i}", generate.serving_lp(generate.seed_of(3, 10_000 + i))))
    out.append(("config1@0.1", generate.config1(scale=0.1)))
    if not quick:
        out.append(("config1", generate.config1()))
        out.append(("config2@0.1", generate.config2(scale=0.1)))
        out.append(("config4-large", generate.serving_lp(generate.seed_of(4, 10_000), 200_000, 1_000_000,
                                                        zipf_rows=1.5)))
    return out


def structural(p) -> ScaledProblem:
    """A load() input with the LP's structure (timing does not depend on the
    values, so the host scaling is skipped)."""
    at = p.a.transpose()
    m, n = p.nrows, p.ncols
    return ScaledProblem(a=p.a, at=at, c=np.zeros(n), lc=np.zeros(m), uc=np.zeros(m), lv=np.zeros(n),
                         uv=np.zeros(n), row_scale=np.ones(m), col_scale=np.ones(n), bound_scale=1.0, obj_scale=1.0)


def sweep(b: HipBackend, items, reps: int = 30, log=print):
    rows = []
    for name, p in items:
        sp = structural(p)
        for side, mat in (("A", sp.a), ("At", sp.at)):
            nnz = int(mat.row_ptr[-1])
            feats = {"matrix": name, "side": side, "rows": int(mat.nrows), "nnz": nnz,
                     "size_class": size_class(nnz), "mean_class": mean_class(nnz / max(1, mat.nrows)),
                     "max_row": int(np.diff(mat.row_ptr).max()) if mat.nrows else 0}
            times = {}
            for cand in CANDIDATES:
                if side == "A":
                    pg = b.load(sp, tuning_a=cand, tuning_at=DEFAULT)
                else:
                    pg = b.load(sp, tuning_a=DEFAULT, tuning_at=cand)
                times[str(cand)] = b.spmv_time(pg, transpose=(side == "At"), reps=reps)
                pg.free()
            feats["ms"] = times
            best = min(times, key=times.get)
            log(f"{name:>14} {side:>2} rows={feats['rows']:>8} nnz={nnz:>9}: default {times[str(DEFAULT)]:.4f} ms, "
                f"best {best} {times[best]:.4f} ms")
            rows.append(feats)
    return rows


REGRET = 1.03  # a cell's plan may be at most 3% slower than the default on any of its matrices


def rule(rows):
    """Per (size class, mean class): among the candidates that are within
    REGRET of the default on every matrix of the cell, the one with the least
    total time (the default when no other qualifies)."""
    cells = {}
    for r in rows:
        cells.setdefault((r["size_class"], r["mean_class"]), []).append(r)
    table = []
    for (sc, mc), rs in sorted(cells.items()):
        ok = [c for c in CANDIDATES
              if all(r["ms"][str(c)] <= REGRET * r["ms"][str(DEFAULT)] for r in rs)]
        tot = {c: sum(r["ms"][str(c)] for r in rs) for c in ok}
        best = min(ok, key=lambda c: (tot[c], c != DEFAULT))
        table.append((sc, mc) + best)
    return table


def write_inc(table, path):
    with open(path, "w") as f:
        f.write("// Generated by saddlepoint_amd/autotune.py; {size_class, mean_class, block_nnz, long_row, max_rows}.\n")
        for t in table:
            f.write("{%d, %d, %d, %d, %d},\n" % t)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--from-sweep", help="rebuild the table from a saved autotune_sweep.json (no GPU)")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--reps", type=int, default=30)
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    if args.from_sweep:
        with open(args.from_sweep) as f:
            rows = json.load(f)["rows"]
        table = rule(rows)
        write_inc(table, os.path.join(args.out, "tuning_rule.inc"))
        print(f"{len(table)} cells written", flush=True)
        return 0
    t0 = time.time()
    items = training_set(args.quick)
    print(f"training set: {len(items)} LPs generated in {time.time() - t0:.1f}s", flush=True)
    b = HipBackend(0)
    rows = sweep(b, items, args.reps, log=lambda s: print(s, flush=True))
    table = rule(rows)
    write_inc(table, os.path.join(args.out, "tuning_rule.inc"))
    with open(os.path.join(args.out, "autotune_sweep.json"), "w") as f:
        json.dump({"candidates": CANDIDATES, "default": DEFAULT, "rows": rows, "table": table}, f)
    print(f"{len(table)} cells written", flush=True)
    b.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
