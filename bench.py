"""Benchmark: restarted Halpern PDHG iterations per second on one MI355X.

Workload: the plain-FP64 single solve at relaxation scale (400k rows × 900k
columns, ~8M nonzeros, Zipf(1.3) column counts, 20 dense rows of 20k-60k
nonzeros; seeded synthetic data); ``--workload config2`` measures the large
instance instead (4M × 10M, ~100M nonzeros). A step is one PDHG iteration of
the full driver. The warm-up and the timed window are whole 64-iteration
check windows (the requested counts rounded up; ``steps`` in the line is the
number actually timed) and the window starts on a check boundary, so the
rate includes the KKT/restart/infeasibility check every 64 iterations; the
problem is resident in HBM when the window starts.

Also reported, on one JSON line:
  roofline     the dominant kernel's algorithmic bytes per launch over its
               average duration (two HIP events around 20 back-to-back
               launches on the solver's stream), against the 8 TB/s HBM
               peak; ``gather`` puts its gathers per second against the
               replay of the same matrix's column indices with nothing but
               the gathers (the ceiling its access pattern allows);
  serving      a fixed config-3 sample (the first 16 LPs of the stream)
               solved one at a time in FP64 and served in FP64 and mixed
               precision: LPs/s, p50/p99 latency;
  cpu_baseline the oracle (a C restatement of the reference's CPU backend and
               driver) on the host cores, over a bounded sample of the same
               workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LPs solved/s + p50/p99 latency at 1e-6 KKT vs plain FP64; iters/s, HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(m: int, n: int, nnz: int) -> dict:
    """Algorithmic HBM bytes per launch of the fused iteration kernels (FP64
    values and vectors, int32 indices; each gathered vector counted once)."""
    return {
        # Aᵀ side: values+indices, row pointers, gather y, read x c lv uv x_anchor, write x_bar x
        "sp_spmv_t_primal_halpern_f64": 12 * nnz + 4 * (n + 1) + 8 * m + 40 * n + 16 * n,
        # A side: values+indices, row pointers, gather x_bar, read y lc uc y_anchor, write y
        "sp_spmv_dual_halpern_f64": 12 * nnz + 4 * (m + 1) + 8 * n + 32 * m + 8 * m,
        "sp_spmv_f64": 12 * nnz + 4 * (m + 1) + 8 * n + 8 * m,
        # Mixed precision (FP32 values and displacements, FP64 anchors/bounds/c and anchor products):
        # Aᵀ side: values+indices 8·nnz, row pointers, gather dy, read dx xa c lv uv atya, write dx_bar dx
        "sp_spmv_t_primal_mixed": 8 * nnz + 4 * (n + 1) + 4 * m + (4 + 8 * 5) * n + 8 * n,
        # A side: values+indices, row pointers, gather dx_bar, read dy ya lc uc axa, write dy
        "sp_spmv_dual_mixed": 8 * nnz + 4 * (m + 1) + 4 * n + (4 + 8 * 4) * m + 4 * m,
    }


def read_traffic(path: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the factor
    the same counter reads on a copy kernel of known size), or None."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 2 and parts[0] == kernel + ":":
                    return float(parts[1])
    except OSError:
        return None
    return None


def cpu_baseline(problem, seconds_budget: float) -> dict:
    from oracle import oracle as O

    cores = min(os.cpu_count() or 1, 16)
    O.set_threads(cores)
    params = O.default_params(tolerance=0.0)
    # One 64-iteration window first, to size five runs totalling ~seconds_budget.
    t64, scale_s = O.time_iterations(problem, 64, params)
    per_iteration = max(t64, 1e-9) / 64
    k = int(seconds_budget / 5 / per_iteration) // 64 * 64
    k = min(max(k, 64), 64 * 32)
    rates = []
    for _ in range(5):
        t, _ = O.time_iterations(problem, k, params)
        rates.append(k / t)
    # A per-core figure: the same restatement on one thread, 64 iterations.
    O.set_threads(1)
    t1, _ = O.time_iterations(problem, 64, params)
    O.set_threads(cores)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": float(np.median(rates)), "unit": "iters/s", "cores": cores, "kind": "port",
            "sample": f"median of 5 runs of {k} iterations of the same instance (check work included, power "
                      f"iteration and {scale_s:.1f}s of host scaling excluded), oracle C restatement with "
                      f"{cores} OpenMP threads",
            "runs": [round(r, 2) for r in rates], "single_thread_value": 64 / t1, "cpu_model": model,
            "nproc": os.cpu_count()}


def pct(v, q):
    return float(np.percentile(np.asarray(v, float), q)) if len(v) else None


def fifo_latencies(arrivals, service):
    """Completion - arrival of a one-at-a-time server taking jobs in arrival
    order with the given service times (exact for a deterministic FIFO queue)."""
    free, out = 0.0, []
    for a, s in zip(arrivals, service):
        start = max(a, free)
        free = start + s
        out.append(free - a)
    return out


def run_served(b, lps, params_of, arrivals=None, max_active=256, with_solution=False, use_graphs=None,
               chunk_work=None, max_skipped=None):
    """Serves `lps` (all at t = 0, or at the given arrival times on the
    server's clock); returns (makespan, latencies, results)."""
    from saddlepoint_amd.serving import Server

    srv = Server(b, max_active=max_active, use_graphs=use_graphs, chunk_work=chunk_work, max_skipped=max_skipped)
    n = len(lps)
    ids = [None] * n
    t_arr = [0.0] * n
    nxt = 0
    t0 = srv.clock()
    if arrivals is None:
        for i, p in enumerate(lps):
            ids[i] = srv.submit(p, params_of(i))
            t_arr[i] = t0
        nxt = n
    done, last_report = 0, time.time()
    while done < n:
        now = srv.clock()
        while nxt < n and t0 + arrivals[nxt] <= now:
            ids[nxt] = srv.submit(lps[nxt], params_of(nxt))
            t_arr[nxt] = t0 + arrivals[nxt]
            nxt += 1
        active, waiting, done = srv.run_chunk()
        if active == 0 and waiting == 0 and nxt < n and done < n:
            time.sleep(max(0.0, t0 + arrivals[nxt] - srv.clock()))
        if time.time() - last_report > 20:
            log(f"  served {done}/{n} (active {active}, waiting {waiting})")
            last_report = time.time()
    t_end = srv.clock()
    res = [srv.result(j, with_solution=with_solution) for j in ids]
    lat = [r.completed - ta for r, ta in zip(res, t_arr)]
    # Saturated throughput: completions while LPs were still waiting for a
    # slot (up to the last admission), per second of that window.
    last_admit = max(r.admitted for r in res)
    window = last_admit - t0
    done_in_window = sum(1 for r in res if r.completed <= last_admit)
    run_served.saturated = (done_in_window / window) if window > 0 and done_in_window > 0 else None
    run_served.window = (window, [i for i, r in enumerate(res) if r.completed <= last_admit])
    srv.close()
    return t_end - t0, lat, [r.result for r in res]


def serving_bench(args) -> dict:
    from saddlepoint_amd import generate
    from saddlepoint_amd.backend import HipBackend, default_params, solve

    t0 = time.time()
    if args.small:
        # Small LPs only (rows log-uniform 1k-5k), from their own seeds.
        lps = [generate.serving_lp(generate.seed_of(args.config, 5000 + i), m_lo=1000, m_hi=5000)
               for i in range(args.lps)]
    else:
        lps = generate.serving_stream(args.config, args.lps)
    tols = [1e-4 if i % 2 == 0 else 1e-6 for i in range(len(lps))]
    nnz = [p.a.nnz for p in lps]
    log(f"generated {len(lps)} LPs of the config-{args.config} stream in {time.time() - t0:.1f}s; "
        f"nnz mean {np.mean(nnz):.0f}, max {max(nnz)}")
    b = HipBackend(0)
    limit = args.lp_time_limit if args.lp_time_limit > 0 else -1.0
    it_cap = args.lp_max_iterations
    autotune = int(args.config == 4)  # config 4: each LP's SpMV plans from the autotuned rule
    modes = [m.strip() for m in args.modes.split(",") if m.strip()]
    if "one" not in modes:
        # Served modes only (for profiling): no baseline, no ratios.
        out = {"lps": len(lps), "config": args.config}
        for prec in [m for m in modes if m in ("fp64", "mixed")]:
            pr = int(prec == "mixed")
            pm = (lambda i: default_params(tolerance=tols[i], precision=pr, time_limit_s=limit,
                                           max_iterations=it_cap, spmv_autotune=autotune * pr,
                                           mixed_stall_checks=args.stall_checks))
            mk, lat, res = run_served(b, lps, pm, max_active=args.max_active, chunk_work=args.chunk_work,
                                      use_graphs=not args.no_graphs, max_skipped=args.max_skipped)
            out["served_" + prec] = {"lps_per_s": len(lps) / mk, "makespan_s": mk, "p50_s": pct(lat, 50),
                                     "p99_s": pct(lat, 99)}
            log(f"served {prec}: {len(lps) / mk:.3f} LPs/s, makespan {mk:.1f}s")
        b.close()
        return out
    # Plain FP64 one at a time: every LP's whole solve (validate, scale, load,
    # solve, download), back to back.
    service, one = [], []
    for i, (p, tol) in enumerate(zip(lps, tols)):
        s = time.perf_counter()
        r = solve(b, p, default_params(tolerance=tol, time_limit_s=limit, max_iterations=it_cap))
        service.append(time.perf_counter() - s)
        one.append(r)
        log(f"  one-at-a-time {i}: m={p.nrows} nnz={p.a.nnz} tol={tol:g} {r.status} {r.iterations} it "
            f"{service[-1]:.3f}s")
    total = sum(service)
    cap = len(lps) / total
    sat_one = fifo_latencies([0.0] * len(lps), service)
    out = {"lps": len(lps), "config": args.config, "tolerances": "50/50 1e-4 / 1e-6",
           "nnz_mean": float(np.mean(nnz)), "nnz_max": int(max(nnz)),
           "one_at_a_time_fp64": {"lps_per_s": cap, "makespan_s": total, "p50_s": pct(sat_one, 50),
                                  "p99_s": pct(sat_one, 99)}}
    rng = np.random.default_rng(12345)
    arrivals = list(np.cumsum(rng.exponential(1.0 / (args.rate_frac * cap), len(lps))))
    arrivals = [a - arrivals[0] for a in arrivals]
    poi_one = fifo_latencies(arrivals, service)
    out["one_at_a_time_fp64"]["poisson"] = {"rate_lps_per_s": args.rate_frac * cap, "p50_s": pct(poi_one, 50),
                                            "p99_s": pct(poi_one, 99)}
    for prec in ("fp64", "mixed"):
        pr = int(prec == "mixed")
        pm = (lambda i: default_params(tolerance=tols[i], precision=pr, time_limit_s=limit, max_iterations=it_cap,
                                       spmv_autotune=autotune * pr, mixed_stall_checks=args.stall_checks))
        mk, lat, res = run_served(b, lps, pm, with_solution=args.check, max_active=args.max_active,
                                  chunk_work=args.chunk_work, max_skipped=args.max_skipped)
        st = {}
        for r in res:
            st[r.status] = st.get(r.status, 0) + 1
        entry = {"lps_per_s": len(lps) / mk, "makespan_s": mk, "p50_s": pct(lat, 50), "p99_s": pct(lat, 99),
                 "statuses": st, "speedup_vs_one_at_a_time": (len(lps) / mk) / cap,
                 "lps_per_s_saturated": run_served.saturated}
        # The same window measured in work: the one-at-a-time service time of
        # the LPs finished in it, per second of it (LPs still running count for
        # nothing, so this underestimates the steady-state speedup).
        window, finished = run_served.window
        if window > 0 and finished:
            entry["work_speedup_saturated"] = sum(service[i] for i in finished) / window
        if args.check:
            # The KKT checker on every served solution (sp_check, on the device).
            from saddlepoint_amd.backend import check
            kk = [check(b, p, r.x, r.y).max_relative() for p, r in zip(lps, res) if r.status == "Optimal"]
            ok = sum(1 for p, r, t in zip(lps, res, tols)
                     if r.status == "Optimal" and check(b, p, r.x, r.y).max_relative() <= 2 * t)
            entry["checked"] = {"optimal": len(kk), "pass_at_2tol": ok, "max_relative_kkt": max(kk) if kk else None}
            for r in res:
                r.x = r.y = None
        if pr:
            gaps = [abs(r.objective - o.objective) / (1 + abs(o.objective)) for r, o in zip(res, one)
                    if r.status == "Optimal" and o.status == "Optimal"]
            kkt = [max(r.primal_residual, r.dual_residual, r.gap) for r in res if r.status == "Optimal"]
            entry["objective_gap_vs_fp64"] = {"max": max(gaps) if gaps else None, "median": pct(gaps, 50)}
            entry["fp64_fallback_lps"] = sum(1 for r in res if r.mixed_fp64_iterations > 0)
            entry["per_instance"] = [{"i": i, "status": r.status, "iterations": r.iterations,
                                      "fp64_iterations": r.iterations if o is None else o.iterations,
                                      "objective_gap": abs(r.objective - o.objective) / (1 + abs(o.objective)),
                                      "primal_res": r.primal_residual, "dual_res": r.dual_residual, "gap": r.gap}
                                     for i, (r, o) in enumerate(zip(res, one))]
        sat = entry["lps_per_s_saturated"]
        log(f"  saturated: {sat if sat is None else round(sat, 3)} LPs/s while LPs waited for a slot; "
            f"work speedup {entry.get('work_speedup_saturated')}")
        log(f"served {prec}: {entry['lps_per_s']:.3f} LPs/s (x{entry['speedup_vs_one_at_a_time']:.2f}), "
            f"p50 {entry['p50_s']:.3f}s p99 {entry['p99_s']:.3f}s")
        if args.poisson and pr:
            _, plat, _ = run_served(b, lps, pm, arrivals, max_active=args.max_active, chunk_work=args.chunk_work,
                                    max_skipped=args.max_skipped)
            entry["poisson"] = {"rate_lps_per_s": args.rate_frac * cap, "p50_s": pct(plat, 50),
                                "p99_s": pct(plat, 99)}
            log(f"  poisson at {args.rate_frac} x capacity: p50 {pct(plat, 50):.3f}s p99 {pct(plat, 99):.3f}s "
                f"(one at a time: p50 {pct(poi_one, 50):.3f}s p99 {pct(poi_one, 99):.3f}s)")
        out["served_" + prec] = entry
    b.close()
    return out


def stream_bench(args) -> dict:
    """Steady state of a serving stream over a window of `--stream` seconds
    per mode, on the same sequence of the config-3 (or, with `--config 4`,
    config-4) stream (instance 0, 1, 2, ...), every LP capped at
    `--lp-max-iterations` in every mode.

    Arrivals are either a closed loop (`--stream-rate 0`: `--stream-backlog`
    LPs are in the system at all times, the next entering when one completes,
    so work is always waiting: the mode's capacity) or Poisson at
    `--stream-rate` LPs/s (seeded exponential gaps, the same arrival times in
    every mode). Completions per second are counted over the last three
    quarters of the window (the first quarter is warm-up) and latencies
    (completion - arrival) taken of the LPs completed in them; LPs that have
    arrived but not completed at the window's end are reported, since an
    overloaded mode's latencies of completed LPs leave them out. One at a time
    solves the sequence back to back in plain FP64, in arrival order."""
    from saddlepoint_amd import generate
    from saddlepoint_amd.backend import HipBackend, default_params, solve
    from saddlepoint_amd.serving import DONE, FAILED, Server

    T, K, rate = args.stream, args.stream_backlog, args.stream_rate
    cfg, it_cap = args.config, args.lp_max_iterations
    t0 = time.time()
    pool = []
    for start_i in range(0, args.stream_pool, 50):
        pool += generate.serving_stream(cfg, min(50, args.stream_pool - start_i), start=start_i)
        log(f"  generated {len(pool)} LPs")
    tols = [1e-4 if i % 2 == 0 else 1e-6 for i in range(len(pool))]
    arrivals = (np.cumsum(np.random.default_rng(2024).exponential(1.0 / rate, len(pool))) if rate > 0 else None)
    log(f"stream: {len(pool)} config-{cfg} LPs generated in {time.time() - t0:.1f}s; window {T}s, "
        + (f"Poisson arrivals at {rate} LPs/s" if rate > 0 else f"closed loop of {K}"))
    b = HipBackend(0)
    workload = ("config-3 stream (rows log-uniform 1k-200k, columns 1.5-3x rows, 4-30 nonzeros per column, "
                "tolerance 50/50 1e-4 / 1e-6)")
    if cfg == 4:
        workload = ("config-4 stream (the config-3 stream with 10% large LPs: rows log-uniform 200k-1M, "
                    "Zipf(1.5) row lengths; served mixed precision with autotuned SpMV plans)")
    out = {"window_s": T, "warmup_s": T / 4, "pool": len(pool), "lp_max_iterations": it_cap,
           "arrivals": f"poisson {rate} LPs/s (seed 2024)" if rate > 0 else f"closed loop, {K} in the system",
           "workload": workload + ", instances taken in order"}
    autotune = int(cfg == 4)  # config 4: served mixed precision takes each LP's SpMV plans from the autotuned rule

    def summary(done, lat, taken, arrived, open_at_end):
        return {"lps_per_s": done / (0.75 * T), "completed_in_window": done, "taken": taken,
                "arrived_by_end": arrived, "in_system_at_end": open_at_end, "pool_exhausted": taken >= len(pool),
                "p50_s": pct(lat, 50), "p99_s": pct(lat, 99)}

    def arrived_by(t):
        return int(np.searchsorted(arrivals, t, side="right")) if rate > 0 else None

    modes = [m.strip() for m in args.modes.split(",") if m.strip()]
    out["server"] = {"chunk_work": args.chunk_work, "max_skipped": args.max_skipped, "max_active": args.max_active}
    # One at a time. Closed loop: LP i enters when LP i - K completes (at the
    # start for the first K). Poisson: LP i enters at its arrival time; the
    # GPU idles while nothing has arrived.
    start = time.perf_counter()
    done, lat, i, ends, last_report = 0, [], 0, [], time.time()
    while i < len(pool) and "one" in modes:
        now = time.perf_counter() - start
        if now >= T:
            break
        if rate > 0 and arrivals[i] > now:
            time.sleep(min(arrivals[i], T) - now)
            continue
        solve(b, pool[i], default_params(tolerance=tols[i], max_iterations=it_cap))
        s1 = time.perf_counter() - start
        if time.time() - last_report > 30:
            log(f"  one at a time: {s1:.0f}s, {i + 1} solved")
            last_report = time.time()
        ends.append(s1)
        if T / 4 <= s1 < T:
            done += 1
            entered = arrivals[i] if rate > 0 else (ends[i - K] if i >= K else 0.0)
            lat.append(s1 - entered)
        i += 1
    if "one" in modes:
        arrived = arrived_by(T) if rate > 0 else i + K
        out["one_at_a_time_fp64"] = summary(done, lat, i, arrived, max(arrived - i, 0) if rate > 0 else K)
        log(f"  one at a time: {out['one_at_a_time_fp64']['lps_per_s']:.3f} LPs/s ({i} taken)")
    for prec in [m for m in ("fp64", "mixed") if m in modes]:
        srv = Server(b, chunk_work=args.chunk_work or None, max_skipped=args.max_skipped or None,
                     max_active=args.max_active)
        pr = int(prec == "mixed")
        entered, i = {}, 0
        start = srv.clock()
        done, lat, last_report = 0, [], time.time()
        pending = set()
        while True:
            now = srv.clock() - start
            if now >= T:
                break
            while i < len(pool) and (len(pending) < K if rate <= 0 else arrivals[i] <= now):
                j = srv.submit(pool[i], default_params(tolerance=tols[i], precision=pr, max_iterations=it_cap,
                                                       spmv_autotune=autotune * pr))
                entered[j] = start + (arrivals[i] if rate > 0 else now)
                pending.add(j)
                i += 1
            if pending:
                srv.run_chunk()
            elif i < len(pool) and rate > 0:
                time.sleep(max(min(arrivals[i], T) - now, 0.0))
                continue
            else:
                break
            for j in [j for j in pending if srv.state(j) in (DONE, FAILED)]:
                pending.discard(j)
                r = srv.result(j, with_solution=False)
                t_rel = r.completed - start
                if T / 4 <= t_rel < T:
                    done += 1
                    lat.append(r.completed - entered[j])
                srv.forget(j)
            if time.time() - last_report > 30:
                log(f"  served {prec}: {srv.clock() - start:.0f}s, {i} taken, {len(pending)} in the server")
                last_report = time.time()
        srv.close()
        arrived = arrived_by(T) if rate > 0 else i
        key = "served_" + prec
        out[key] = summary(done, lat, i, arrived, len(pending) + (arrived - i if rate > 0 else 0))
        if "one_at_a_time_fp64" in out:
            out[key]["speedup_vs_one_at_a_time"] = out[key]["lps_per_s"] / max(
                out["one_at_a_time_fp64"]["lps_per_s"], 1e-12)
        log(f"  served {prec}: {out[key]['lps_per_s']:.3f} LPs/s ({i} taken)")
    b.close()
    return out


def large_bench(args) -> dict:
    """The large single instance (config 2), plain FP64 against the mixed
    precision mode. Both run the same iteration cap (`--max-iterations`, no
    time limit unless `--time-limit` > 0) at 1e-4, so a mode that does not
    converge stops at the same iteration count as the other and the two are
    compared at equal work: objective gap, distance of the iterates, and each
    one's KKT report from the device checker. A mode that reaches 1e-4 within
    the cap is run again at 1e-6 with the same cap."""
    from saddlepoint_amd import generate
    from saddlepoint_amd.backend import HipBackend, check, default_params, solve

    t0 = time.time()
    p = generate.config2(scale=args.scale)
    log(f"generated config2 m={p.nrows} n={p.ncols} nnz={p.a.nnz} in {time.time() - t0:.1f}s")
    b = HipBackend(0)
    out = {"m": p.nrows, "n": p.ncols, "nnz": p.a.nnz, "scale": args.scale, "max_iterations": args.max_iterations,
           "runs": []}
    limit = args.time_limit if args.time_limit > 0 else -1.0
    tols = [1e-4]
    for tol in (1e-4, 1e-6):
        if tol not in tols:
            break
        res = {}
        for prec in [m for m in ("fp64", "mixed") if m in args.modes.split(",")]:
            prm = default_params(tolerance=tol, precision=int(prec == "mixed"), max_iterations=args.max_iterations,
                                 time_limit_s=limit, log_every_checks=args.log_every)
            t = time.perf_counter()
            r = solve(b, p, prm)
            wall = time.perf_counter() - t
            k = check(b, p, r.x, r.y)
            res[prec] = r
            e = {"tolerance": tol, "precision": prec, "status": r.status, "iterations": r.iterations,
                 "restarts": r.restarts, "setup_s": r.setup_time, "solve_s": r.solve_time, "wall_s": wall,
                 "iters_per_s": r.iterations / max(r.solve_time, 1e-9), "objective": r.objective,
                 "kkt_relative_primal": k.relative_primal, "kkt_relative_dual": k.relative_dual,
                 "kkt_relative_gap": k.relative_gap, "fp64_iterations_in_mixed": r.mixed_fp64_iterations}
            if prec == "mixed" and "fp64" in res:
                f = res["fp64"]
                e["equal_iterations"] = r.iterations == f.iterations
                e["objective_gap_vs_fp64"] = abs(r.objective - f.objective) / (1 + abs(f.objective))
                e["x_rel_diff_vs_fp64"] = float(np.max(np.abs(r.x - f.x)) / (1 + np.max(np.abs(f.x))))
                e["y_rel_diff_vs_fp64"] = float(np.max(np.abs(r.y - f.y)) / (1 + np.max(np.abs(f.y))))
            log(f"  tol {tol:g} {prec}: {r.status} {r.iterations} it, solve {r.solve_time:.2f}s "
                f"({e['iters_per_s']:.0f} it/s), setup {r.setup_time:.2f}s, kkt max {k.max_relative():.2e}")
            out["runs"].append(e)
        if any(r.status == "Optimal" for r in res.values()):
            tols.append(1e-6)
    if "fp64" in args.modes.split(",") and "mixed" in args.modes.split(","):
        # Host scaling of the same LP, for the setup comparison.
        t = time.perf_counter()
        from saddlepoint_amd.backend import ScaledProblem
        ScaledProblem.new(p)
        out["host_scaling_s"] = time.perf_counter() - t
        log(f"  host ScaledProblem::new: {out['host_scaling_s']:.2f}s")
    b.close()
    return out


def share_bench(args) -> dict:
    """One GPU's share of config 2 split across P GPUs by row blocks: A' in P
    contiguous row blocks and A'ᵀ in P contiguous row blocks, each cut where
    the nonzero count crosses a multiple of nnz/P. Every block is loaded on
    this GPU at its real shape (all columns, so the gathered table is the
    full x̄ or y) and its SpMV timed; the slowest block of each side is the
    per-iteration critical path, beside the all-gathers such a split needs
    (y and x̄ in full, 8(m+n) bytes, of which (P-1)/P arrive over xGMI)."""
    from saddlepoint_amd import generate
    from saddlepoint_amd.autotune import structural
    from saddlepoint_amd.backend import HipBackend
    from saddlepoint_amd.problem import CsrMatrix, LpProblem

    P = args.share
    t0 = time.time()
    p = generate.config2(scale=args.scale)
    at = p.a.transpose()
    log(f"generated config2 m={p.nrows} n={p.ncols} nnz={p.a.nnz} in {time.time() - t0:.1f}s")
    b = HipBackend(0)

    def lp_of(a):
        z = np.zeros
        return LpProblem(a=a, c=z(a.ncols), lc=z(a.nrows), uc=z(a.nrows), lv=z(a.ncols), uv=z(a.ncols))

    def timed(a, reps=20):
        pg = b.load(structural(lp_of(a)))
        try:
            return b.spmv_time(pg, transpose=False, reps=reps)
        finally:
            pg.free()

    out = {"m": p.nrows, "n": p.ncols, "nnz": p.a.nnz, "gpus": P, "scale": args.scale}
    for side, mat in (("A", p.a), ("At", at)):
        rp = mat.row_ptr
        cuts = np.searchsorted(rp, np.arange(P + 1) * (rp[-1] / P), side="left")
        cuts[0], cuts[-1] = 0, mat.nrows
        blocks = []
        for r0, r1 in zip(cuts[:-1], cuts[1:]):
            k0, k1 = int(rp[r0]), int(rp[r1])
            blk = CsrMatrix(int(r1 - r0), mat.ncols, (rp[r0:r1 + 1] - rp[r0]).copy(), mat.col_idx[k0:k1].copy(),
                            mat.values[k0:k1].copy())
            ms = timed(blk)
            blocks.append({"rows": int(r1 - r0), "nnz": k1 - k0, "us": ms * 1e3})
            log(f"  {side} block rows {r0}-{r1}: nnz {k1 - k0}, {ms * 1e3:.1f} us")
        full = timed(mat, reps=10)
        log(f"  {side} whole on one GPU: {full * 1e3:.1f} us")
        out[side] = {"blocks": blocks, "max_block_us": max(x["us"] for x in blocks), "one_gpu_us": full * 1e3}
    gather_bytes = 8 * (p.nrows + p.ncols) * (P - 1) / P
    out["allgather_bytes_per_gpu"] = gather_bytes
    out["spmv_critical_us"] = out["A"]["max_block_us"] + out["At"]["max_block_us"]
    out["spmv_one_gpu_us"] = out["A"]["one_gpu_us"] + out["At"]["one_gpu_us"]
    b.close()
    return out


# The reference's published GPU rows (bench/results/2026-11-03-rtx-a1000-pid.csv; RTX A1000 Laptop GPU,
# MIPLIB 2017 LP relaxations at 1e-4; the two rows that hit the 300 s limit are left out): instance, rows,
# cols, nnz, iterations, solve seconds.
REFERENCE_GPU_ROWS = [
    ("rail507", 509, 63_019, 468_878, 3_456, 0.758),
    ("irish-electricity", 104_259, 61_728, 523_257, 86_592, 26.416),
    ("supportcase10", 165_684, 14_770, 555_082, 19_264, 6.681),
    ("ex10", 69_608, 17_680, 1_162_000, 448, 0.272),
    ("s100", 14_733, 364_417, 1_777_917, 174_848, 176.365),
    ("savsched1", 295_989, 328_575, 1_770_507, 704, 0.914),
    ("neos-5052403-cygnet", 38_268, 32_868, 4_898_304, 22_720, 28.627),
]


def reference_shapes_bench(args) -> dict:
    """Per-iteration time on random LPs of the published instances' sizes
    (the instance files are not available here), beside the reference's
    published GPU time per iteration (solve seconds / iterations, power
    iteration and checks included). Ours: `steps` iterations after `warmup`,
    checks every 64 included, power iteration excluded."""
    from saddlepoint_amd import generate
    from saddlepoint_amd.backend import HipBackend, Solver, default_params

    b = HipBackend(0)
    rows = []
    for i, (name, m, n, nnz, its, secs) in enumerate(REFERENCE_GPU_ROWS):
        p = generate.shaped_lp(generate.seed_of(9, i), m, n, nnz)
        s = Solver(b, p, default_params(tolerance=0.0, fused=1))
        s.advance(64)
        b.sync()
        t = time.perf_counter()
        s.advance(args.steps)
        b.sync()
        el = time.perf_counter() - t
        s.close()
        ours = 1e6 * el / args.steps
        ref = 1e6 * secs / its
        rows.append({"instance": name, "m": m, "n": n, "nnz_published": nnz, "nnz": int(p.a.nnz),
                     "reference_us_per_iteration": round(ref, 1), "us_per_iteration": round(ours, 2),
                     "ratio": round(ref / ours, 1)})
        log(f"  {name:>20}: {m}x{n}, {p.a.nnz} nnz: {ours:.1f} us/iteration (reference GPU {ref:.1f}, x{ref / ours:.1f})")
    b.close()
    return {"rows": rows, "steps": args.steps}


def full_solve_bench(args) -> dict:
    """Config 1 solved to 1e-6 on the GPU (plain FP64, the default options),
    its KKT measures from the device checker."""
    from saddlepoint_amd import generate
    from saddlepoint_amd.backend import HipBackend, check, default_params, solve

    p = generate.config1(scale=args.scale)
    log(f"generated config1 m={p.nrows} n={p.ncols} nnz={p.a.nnz}; progress lines every 200 checks (solver.rs:367-378)")
    b = HipBackend(0)
    prm = default_params(tolerance=1e-6, max_iterations=args.max_iterations,
                         time_limit_s=args.time_limit if args.time_limit > 0 else -1.0,
                         log_every_checks=200)
    t = time.perf_counter()
    r = solve(b, p, prm)
    wall = time.perf_counter() - t
    k = check(b, p, r.x, r.y)
    out = {"m": p.nrows, "n": p.ncols, "nnz": p.a.nnz, "gpu": {
        "status": r.status, "iterations": r.iterations, "restarts": r.restarts, "objective": r.objective,
        "solve_s": r.solve_time, "setup_s": r.setup_time, "wall_s": wall,
        "iters_per_s": r.iterations / max(r.solve_time, 1e-9), "kkt_relative_primal": k.relative_primal,
        "kkt_relative_dual": k.relative_dual, "kkt_relative_gap": k.relative_gap}}
    log(f"  gpu: {r.status} {r.iterations} it, {r.restarts} restarts, solve {r.solve_time:.2f}s, "
        f"objective {r.objective:.10e}, kkt max {k.max_relative():.2e}")
    b.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1280)
    ap.add_argument("--warmup", type=int, default=128)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the workload (tests only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="config1",
                    help="the LP the iteration rate is measured on (config1: the headline)")
    ap.add_argument("--panels", type=int, default=-1,
                    help="sp_params.column_panels for the iteration bench (-1: the library's choice; 0: off)")
    ap.add_argument("--hot", type=int, default=-1, choices=[-1, 0, 1],
                    help="sp_params.hot_table for the iteration bench (-1: the library's choice; 0: off)")
    ap.add_argument("--rows", type=int, default=-1, choices=[-1, 0, 1],
                    help="sp_params.reorder_rows for the iteration bench (-1: the library's choice; 0: off)")
    ap.add_argument("--serving-sample", type=int, default=16,
                    help="LPs of the config-3 stream served beside the headline (0: none)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--probe", choices=["copy", "triad", "gather", "copy_f32", "triad_f32", "overlap"],
                    help="run only a bandwidth probe (copy: 2^27 doubles, the HBM counter calibration)")
    ap.add_argument("--table", type=int, default=900_000, help="gather probe: table length in doubles")
    ap.add_argument("--large", action="store_true",
                    help="config 2 instead: FP64 vs mixed precision at 1e-4 and 1e-6 (use --scale to shrink)")
    ap.add_argument("--max-iterations", type=int, default=200_000)
    ap.add_argument("--log-every", type=int, default=0,
                    help="--large: the solver's progress line every this many checks (0: none)")
    ap.add_argument("--time-limit", type=float, default=0.0, help="--large, --full-solve: seconds per solve (0: none)")
    ap.add_argument("--full-solve", action="store_true",
                    help="config 1 solved to 1e-6 on the GPU (KKT from the device checker)")
    ap.add_argument("--reference-shapes", action="store_true",
                    help="us/iteration on random LPs of the reference's published instance sizes, beside its figures")
    ap.add_argument("--share", type=int, default=0,
                    help="config 2's SpMVs on one GPU's share of a P-way row-block split (P = this value)")
    ap.add_argument("--stream", type=float, default=0.0,
                    help="steady-state serving over a window of this many seconds (saturated, per mode)")
    ap.add_argument("--stream-backlog", type=int, default=64, help="--stream: LPs kept in the system (closed loop)")
    ap.add_argument("--stream-rate", type=float, default=0.0,
                    help="--stream: Poisson arrivals at this many LPs/s instead of the closed loop")
    ap.add_argument("--stream-pool", type=int, default=600, help="--stream: LPs of the sequence generated up front")
    ap.add_argument("--serving", action="store_true",
                    help="serving stream instead: LPs/s and latency, served (FP64 and mixed) vs one at a time")
    ap.add_argument("--config", type=int, default=3, choices=[3, 4],
                    help="serving and --stream: the config-3 stream, or config 4 (10%% large LPs)")
    ap.add_argument("--lps", type=int, default=64, help="serving: LPs in the sample")
    ap.add_argument("--small", action="store_true", help="serving: small LPs only (rows log-uniform 1k-5k)")
    ap.add_argument("--modes", default="one,fp64,mixed",
                    help="serving and --stream: modes to run (without 'one' only the served ones, no ratios)")
    ap.add_argument("--no-graphs", action="store_true", help="serving: eager launches instead of hipGraph replay")
    ap.add_argument("--chunk-work", type=int, default=0,
                    help="serving: sp_server_params.chunk_work (0: the default; negative: every LP every chunk)")
    ap.add_argument("--max-active", type=int, default=256, help="serving and --stream: LPs admitted at once")
    ap.add_argument("--max-skipped", type=int, default=0,
                    help="serving and --stream: sp_server_params.max_skipped (0: the library's default)")
    ap.add_argument("--poisson", action="store_true", help="serving: also Poisson arrivals (mixed precision)")
    ap.add_argument("--rate-frac", type=float, default=0.8, help="serving: arrival rate / one-at-a-time capacity")
    ap.add_argument("--lp-time-limit", type=float, default=0.0, help="serving: seconds per LP (0: none)")
    ap.add_argument("--lp-max-iterations", type=int, default=1_000_000,
                    help="serving: iteration cap per LP, the same in every mode (a multiple of 64)")
    ap.add_argument("--check", action="store_true", help="serving: KKT-check every served solution on the device")
    ap.add_argument("--stall-checks", type=int, default=0,
                    help="serving, mixed precision: sp_params.mixed_stall_checks (0: the library's default)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_config1.txt"),
                    help="measured HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE summary)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (this backend has no "
              "multi-GPU path; LPs shard across GPUs as independent instances)", file=sys.stderr)
        return 2

    from saddlepoint_amd.backend import HipBackend

    if args.large:
        r = large_bench(args)
        print(json.dumps({"metric": METRIC, "large": r}), flush=True)
        return 0

    if args.full_solve:
        r = full_solve_bench(args)
        print(json.dumps({"metric": METRIC, "full_solve": r}), flush=True)
        return 0

    if args.reference_shapes:
        r = reference_shapes_bench(args)
        print(json.dumps({"metric": METRIC, "reference_shapes": r}), flush=True)
        return 0

    if args.share:
        r = share_bench(args)
        print(json.dumps({"metric": METRIC, "share": r}), flush=True)
        return 0

    if args.stream:
        r = stream_bench(args)
        print(json.dumps({"metric": METRIC, "stream": r}), flush=True)
        return 0

    if args.serving:
        r = serving_bench(args)
        if "one_at_a_time_fp64" not in r:
            print(json.dumps({"metric": METRIC, "serving_modes": r}), flush=True)
            return 0
        sm = r["served_mixed"]
        print(json.dumps({"metric": METRIC, "value": round(sm["lps_per_s"], 4), "unit": "LPs/s", "n_gpus": 1,
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
                          "data": "synthetic (seeded serving-stream generator)",
                          "config": {"workload": f"config{args.config} serving stream sample: {args.lps} LPs, rows "
                                                 f"log-uniform " + ("1k-5k" if args.small else "1k-200k") +
                                                 (" plus 10% large (200k-1M rows, Zipf(1.5) "
                                                  "row lengths)" if args.config == 4 and not args.small else "") +
                                                 ", tolerance 50/50 1e-4/1e-6, all submitted at t=0",
                                     "parallelism": "single GPU"},
                          "serving": r}), flush=True)
        return 0

    if args.probe == "overlap":
        # Random gathers beside a coalesced stream (8M of each, the size of a
        # config-1 SpMV): alone, in alternate workgroups, and in every
        # workgroup one after the other, from tables of 128 KB to 80 MB.
        b = HipBackend(0)
        n = 1 << 23
        for table in (16384, 400_000, 900_000, 10_000_000):
            t = {k: b.probe("overlap_" + k, n, table, reps=20) * 1e3
                 for k in ("gather", "stream", "alternate", "both")}
            print(json.dumps({"probe": "overlap", "elements": n, "table_doubles": table,
                              "us_gathers_only": round(t["gather"], 2), "us_stream_only": round(t["stream"], 2),
                              "us_alternate_workgroups": round(t["alternate"], 2),
                              "us_both_in_every_workgroup": round(t["both"], 2),
                              "us_sum_of_alone": round(t["gather"] + t["stream"], 2),
                              "g_gathers_per_s_alone": round(n / t["gather"] / 1e3, 1),
                              "stream_gbs_alone": round(12 * n / t["stream"] / 1e3, 1)}), flush=True)
        b.close()
        return 0
    if args.probe:
        b = HipBackend(0)
        n = 1 << 27
        table = args.table if args.probe == "gather" else 0
        ms = b.probe(args.probe, n, table, reps=10)
        per = {"copy": 16, "triad": 24, "gather": 12, "copy_f32": 8, "triad_f32": 12}[args.probe]
        print(json.dumps({"probe": args.probe, "elements": n, "table": table, "ms_per_launch": ms,
                          "bytes_per_launch": per * n, "gbs": per * n / ms / 1e6}), flush=True)
        return 0

    out = iteration_bench(args)
    print(json.dumps(out), flush=True)
    return 0


CHECK = 64  # the driver's check cadence (solver.rs:340): windows start and end on a check


def whole_windows(k: int) -> int:
    """k rounded up to whole 64-iteration check windows (at least one)."""
    return max(CHECK, -(-int(k) // CHECK) * CHECK)


WORKLOADS = {
    "config1": "config1 plain-FP64 single solve: 400k x 900k LP, ~8M nnz, Zipf(1.3) columns + 20 dense rows "
               "(columns relabeled by popularity, the default)",
    "config2": "config2 large single instance: 4M x 10M LP, ~100M nnz, Zipf(1.5) row lengths, uniform columns "
               "(columns relabeled by popularity, the default)",
}


def iteration_bench(args) -> dict:
    """Iterations per second of the full driver on one LP resident in HBM.

    The warm-up and the timed window are whole 64-iteration check windows
    (the requested counts rounded up), and the timed window starts on a check
    boundary, so the rate includes the check work at its real share: one
    check (KKT, restart residuals, infeasibility rays, PID) per 64
    iterations."""
    from saddlepoint_amd import generate
    from saddlepoint_amd.backend import HipBackend, Solver, default_params

    mixed = args.precision == "mixed"
    t0 = time.time()
    p = generate.config1(scale=args.scale) if args.workload == "config1" else generate.config2(scale=args.scale)
    m, n, nnz = p.nrows, p.ncols, p.a.nnz
    log(f"generated {args.workload} m={m} n={n} nnz={nnz} in {time.time() - t0:.1f}s")

    b = HipBackend(0)
    log(f"backend: {b.name()}")
    t0 = time.time()
    s = Solver(b, p, default_params(tolerance=1e-6, fused=1, precision=int(mixed), column_panels=args.panels,
                                    hot_table=args.hot, reorder_rows=args.rows))
    log(f"solver ready (scale + load + power iteration) in {time.time() - t0:.1f}s")
    warmup, steps = whole_windows(args.warmup), whole_windows(args.steps)
    s.advance(warmup)
    b.sync()
    log(f"warmup {warmup} iterations done (requested {args.warmup})")

    # Timed region: whole check windows, nothing but the solver on the stream.
    it0 = s.result(with_solution=False).iterations
    assert it0 % CHECK == 0, it0
    t0 = time.perf_counter()
    done = s.advance(steps)
    b.sync()
    t1 = time.perf_counter()
    res = s.result(with_solution=False)
    ran = res.iterations - it0
    elapsed = t1 - t0
    log(f"timed {ran} iterations (iterations {it0 + 1}-{it0 + ran}, {ran // CHECK} checks; requested {args.steps}) "
        f"in {elapsed:.3f}s (terminated={done}, status={res.status})")

    # Per-kernel breakdown: the same number of iterations again with a HIP
    # event pair around every launch (adds a little to each kernel).
    b.profile(True)
    s.advance(steps)
    b.sync()
    prof = b.profile_read()
    b.profile(False)
    kb = kernel_bytes(m, n, nnz)
    per_kernel = {}
    for name, (ms, launches) in prof.items():
        if launches:
            per_kernel[name] = {"launches": launches, "avg_us_event_bracketed": round(1e3 * ms / launches, 2),
                                "total_ms": round(ms, 3)}
    sides = {0: "sp_spmv_t_primal_mixed" if mixed else "sp_spmv_t_primal_halpern_f64",
             1: "sp_spmv_dual_mixed" if mixed else "sp_spmv_dual_halpern_f64"}
    if any(k.startswith("sp_spmv_panel") for k in per_kernel):
        # Column panels: the A side is the panel SpMV, the short rows' SpMV and
        # the long rows' finishing pass, timed together.
        label = f"A side ({'sp_spmv_panel_f32' if mixed else 'sp_spmv_panel_f64'} + {sides[1]} + finishing)"
        kb[label] = kb[sides[1]]
        sides[1] = label
    hot = [k for k in per_kernel if k.startswith(sides[0] if mixed else sides[0][:-4]) and k != sides[0]]
    if hot:
        # The A'-transpose side with a hot table of y in LDS (sp_params.hot_table).
        kb[hot[0]] = kb[sides[0]]
        sides[0] = hot[0]
    # Each side's main kernel timed without per-launch events: two events
    # around 20 back-to-back launches on copies of the vectors it writes.
    for side, name in sides.items():
        us = 1e3 * s.kernel_time(side, reps=20)
        e = per_kernel.setdefault(name, {})
        e["avg_us"] = round(us, 2)
        e["gbs"] = round(kb[name] / (us * 1e-6) / 1e9, 1)
        e["bytes_per_launch"] = kb[name]
    dom = max(sides.values(), key=lambda k: per_kernel[k]["avg_us"])
    ach = per_kernel[dom]["gbs"]
    b_iter = 24 * nnz + 52 * m + 68 * n if not mixed else 16 * nnz + 48 * m + 60 * n
    traffic = None
    if args.scale == 1.0:
        traffic = read_traffic(os.path.join(ROOT, "profiles", f"traffic_{args.workload}{'_mixed' if mixed else ''}.txt"),
                               dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_launch": kb[dom]}
    # The kernel's practical ceiling: the same matrix's gathers alone, in the
    # SpMV's order and tiling, from a table as long as the gathered vector
    # (x-bar on the A side, y on the A-transpose side).
    side = 1 if dom == sides[1] else 0
    replay_ms = b.gather_replay_time(s.problem(), transpose=side == 0, fp32=mixed, reps=20)
    kernel_rate = nnz / (per_kernel[dom]["avg_us"] * 1e-6)
    replay_rate = nnz / (replay_ms * 1e-3)
    gathered = n if side == 1 else m
    roofline["gather"] = {"gathers_per_launch": nnz, "table_bytes": (4 if mixed else 8) * gathered,
                          "achieved_per_s": round(kernel_rate / 1e9, 1), "ceiling_per_s": round(replay_rate / 1e9, 1),
                          "ceiling": "the matrix's own column-index stream replayed, gathers only",
                          "unit": "G gathers/s", "frac": round(kernel_rate / replay_rate, 3)}
    other = sides[1 - side]
    other_replay = b.gather_replay_time(s.problem(), transpose=side == 1, fp32=mixed, reps=20)
    roofline["gather_other_side"] = {
        "kernel": other, "achieved_per_s": round(nnz / (per_kernel[other]["avg_us"] * 1e-6) / 1e9, 1),
        "ceiling_per_s": round(nnz / (other_replay * 1e-3) / 1e9, 1),
        "frac": round(other_replay * 1e3 / per_kernel[other]["avg_us"], 3)}

    out = {
        "metric": METRIC,
        "value": round(ran / elapsed, 2),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": ran,
        "warmup": warmup,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "ms_per_step": round(1e3 * elapsed / ran, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic (seeded generator; random instance of the named shape)",
        "config": {"workload": WORKLOADS[args.workload] + ("; mixed precision" if mixed else ""), "m": m, "n": n,
                   "column_panels": args.panels, "hot_table": args.hot, "reorder_rows": args.rows,
                   "nnz": nnz, "scale": args.scale, "timed_window": f"iterations {it0 + 1}-{it0 + ran}, "
                   f"{ran // CHECK} whole check windows (checks included)", "parallelism": "single GPU"},
        "iteration_gbs": round(b_iter * ran / elapsed / 1e9, 1),
        "iteration_roofline_frac": round(b_iter * ran / elapsed / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "kernels": per_kernel,
    }
    s.close()
    if args.serving_sample > 0:
        out["serving"] = serving_sample(b, args.serving_sample)
    b.close()
    if not args.no_cpu and args.workload == "config1":
        log("timing the CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    return out


def serving_sample(b, count: int) -> dict:
    """The headline's serving figures on a fixed sample: the first `count`
    LPs of the config-3 stream (rows log-uniform 1k-200k, tolerance 50/50
    1e-4 / 1e-6), all submitted at t = 0, solved one at a time in plain FP64
    and served (continuous batching) in FP64 and in mixed precision."""
    from saddlepoint_amd import generate
    from saddlepoint_amd.backend import default_params, solve

    t0 = time.time()
    lps = generate.serving_stream(3, count)
    tols = [1e-4 if i % 2 == 0 else 1e-6 for i in range(count)]
    log(f"serving sample: {count} config-3 LPs generated in {time.time() - t0:.1f}s")
    service, one = [], []
    for p, tol in zip(lps, tols):
        t = time.perf_counter()
        one.append(solve(b, p, default_params(tolerance=tol)))
        service.append(time.perf_counter() - t)
    total = sum(service)
    lat_one = fifo_latencies([0.0] * count, service)
    out = {"lps": count, "workload": f"config-3 stream instances 0-{count - 1}, all submitted at t=0",
           "tolerances": "50/50 1e-4 / 1e-6",
           "one_at_a_time_fp64": {"lps_per_s": count / total, "makespan_s": total, "p50_s": pct(lat_one, 50),
                                  "p99_s": pct(lat_one, 99)}}
    log(f"  one at a time: {count / total:.3f} LPs/s, makespan {total:.1f}s")
    for prec in ("fp64", "mixed"):
        pr = int(prec == "mixed")
        mk, lat, res = run_served(b, lps, lambda i: default_params(tolerance=tols[i], precision=pr))
        st = {}
        for r in res:
            st[r.status] = st.get(r.status, 0) + 1
        out["served_" + prec] = {"lps_per_s": count / mk, "makespan_s": mk, "p50_s": pct(lat, 50),
                                 "p99_s": pct(lat, 99), "speedup_vs_one_at_a_time": total / mk, "statuses": st}
        log(f"  served {prec}: {count / mk:.3f} LPs/s (x{total / mk:.2f}), p50 {pct(lat, 50):.2f}s "
            f"p99 {pct(lat, 99):.2f}s")
    return out


if __name__ == "__main__":
    sys.exit(main())
