"""``sp-bench`` for this backend (crates/saddlepoint-cli/src/bin/sp-bench.rs).

    python -m saddlepoint_amd.spbench [--solvers gpu,gpu-mixed] [--tol T] [--time-limit SECONDS]
                                      [--device N] [--termination projected|cupdlpx] [--out FILE.csv] FILES...

Solves each MPS file with each selected solver and writes one CSV row per
(instance, solver) to stdout, and to ``--out`` as it goes (appending, with the
header only when the file is new or empty), in the reference's schema
(sp-bench.rs:22-23):

    instance,rows,cols,nnz,solver,status,iterations,solve_s,total_s,objective,rel_primal,rel_dual,rel_gap

Residuals come from the KKT checker run on the returned solution (here the
device checker, ``sp_check``). Solvers: ``gpu`` (plain FP64) and ``gpu-mixed``
(mixed precision); with ``--termination cupdlpx`` they are labelled ``+step``
as the reference labels its rows. The reference's ``cpu`` and ``cupdlpx``
solvers have no counterpart here.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

from .backend import HipBackend, check, default_params, solve
from .mps import read_mps_file_native

HEADER = "instance,rows,cols,nnz,solver,status,iterations,solve_s,total_s,objective,rel_primal,rel_dual,rel_gap"
SOLVERS = {"gpu": 0, "gpu-mixed": 1}


def row(backend, p, solver: str, args) -> str:
    """``status,iterations,solve_s,total_s,objective,rel_primal,rel_dual,rel_gap`` (sp-bench.rs:150-172)."""
    params = default_params(tolerance=args.tol, time_limit_s=args.time_limit, precision=SOLVERS[solver],
                            termination=1 if args.termination == "cupdlpx" else 0)
    t = time.perf_counter()
    try:
        s = solve(backend, p, params)
    except Exception as e:  # the reference reports Error(...) in the status column
        return f"Error({str(e).replace(',', ';')}),,,,,,,"
    total = time.perf_counter() - t
    r = check(backend, p, s.x, s.y)
    return (f"{s.status},{s.iterations},{s.solve_time:.3f},{total:.3f},{s.objective:.12e},"
            f"{r.relative_primal:.3e},{r.relative_dual:.3e},{r.relative_gap:.3e}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="spbench")
    ap.add_argument("--solvers", default="gpu")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--time-limit", type=float, default=600.0)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--termination", choices=["projected", "cupdlpx"], default="projected")
    ap.add_argument("--out")
    ap.add_argument("files", nargs="+")
    args = ap.parse_args(argv)
    solvers = [s.strip() for s in args.solvers.split(",") if s.strip()]
    for s in solvers:
        if s not in SOLVERS:
            print(f"error: unknown solver {s} (this backend offers {', '.join(SOLVERS)})", file=sys.stderr)
            return 2
    backend = HipBackend(args.device)
    out = None
    if args.out:
        fresh = not os.path.exists(args.out) or os.path.getsize(args.out) == 0
        out = open(args.out, "a")
        if fresh:
            out.write(HEADER + "\n")
            out.flush()
    print(HEADER, flush=True)
    try:
        for f in args.files:
            p = read_mps_file_native(f)
            size = f"{p.name},{p.nrows},{p.ncols},{p.a.nnz}"
            for s in solvers:
                print(f"{p.name} on {s} ...", file=sys.stderr, flush=True)
                label = f"{s}+step" if args.termination == "cupdlpx" else s
                line = f"{size},{label},{row(backend, p, s, args)}"
                print(line, flush=True)
                if out:
                    out.write(line + "\n")
                    out.flush()
    finally:
        if out:
            out.close()
        backend.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
