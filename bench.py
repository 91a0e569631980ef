#!/usr/bin/env python3
"""Benchmark of the row-by-row polyomino tiling search on one MI355X.

Workload (the metric's configuration that fits one GPU): 256 seeded 8x8
boards with 4 blocked cells (every open region's area a multiple of 5), the
12 free pentominoes used once each, all tilings counted.  One step = one
batched solve of all 256 boards (every row of every board, one shared device
frontier).  Reported:

  value            records extended/s (parents expanded, summed over rows)
  solutions_per_s  tilings found per second
  plain            the same boards one CudaSearcher at a time (a sample),
                   records/s, solutions/s and per-board latency
  roofline         the extend kernel's algorithmic HBM bytes / its average
                   launch time (HIP events on the engine's stream), against
                   MI355X's 8 TB/s
  cpu_baseline     the oracle (CPU restatement of the reference backtracker)
                   on a bounded sample of the same boards, P host threads

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "d4-pieces_amd"))

METRIC = ("records extended/s + solutions/s on 1 MI355X, batched vs plain; "
          "p99 latency")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(v, q):
    v = sorted(v)
    if not v:
        return None
    k = (len(v) - 1) * q
    lo, hi = int(k), min(int(k) + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def cpu_baseline(boards, gpu_counts, shapes, seconds, threads):
    """Oracle solve_count on the first boards of the workload until about
    `seconds` of wall time; returns the baseline object and a parity flag."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    oracle = Oracle(path)
    done, sols, ok = 0, 0, True
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while done < len(boards) and time.perf_counter() - t0 < seconds:
            chunk = boards[done:done + threads]
            res = list(ex.map(lambda b: oracle.count(shapes, [1] * 12, b), chunk))
            for j, c in enumerate(res):
                ok &= c == gpu_counts[done + j]
            sols += sum(res)
            done += len(chunk)
            log(f"cpu baseline {done} boards {time.perf_counter() - t0:.1f} s")
    dt = time.perf_counter() - t0
    return {
        "value": sols / dt,
        "unit": "solutions/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {done} of {len(boards)} boards, one board per thread, "
                  f"{dt:.1f} s wall, {done / dt:.3f} boards/s",
    }, ok, done


def serve_run(d4, lib, reqs, slots):
    """Open-loop replay of a request stream (arrival times in seconds from
    the start) through one Server; returns latencies (ms), wall time and the
    results."""
    srv = d4.Server(lib, slots=slots, max_height=8)
    n = len(reqs)
    t_arr, sub, res, lat, last_log = {}, 0, {}, [], -1
    t0 = time.perf_counter()
    while len(res) < n:
        now = time.perf_counter() - t0
        while sub < n and reqs[sub][0] <= now:
            rid = srv.submit(reqs[sub][1], reqs[sub][2])
            t_arr[rid] = reqs[sub][0]
            sub += 1
        fl, wt = srv.step() if (sub > len(res)) else (0, 0)
        while True:
            r = srv.poll()
            if r is None:
                break
            res[r[0]] = (r[1], r[2])
            lat.append((time.perf_counter() - t0 - t_arr[r[0]]) * 1e3)
        if len(res) and len(res) % 200 == 0 and len(res) != last_log:
            last_log = len(res)
            log(f"  served {len(res)}/{n}")
        if sub == len(res) and sub < n:
            # idle until the next arrival
            time.sleep(max(0.0, reqs[sub][0] - (time.perf_counter() - t0)))
    return lat, time.perf_counter() - t0, res, srv.stats()


def serve_section(d4, lib, workloads, rates, n_req, slots):
    """Served requests (Poisson arrivals, 70% first-solution) against the
    same requests solved one per CudaSearcher, one after another."""
    out = {"rates": []}
    base_reqs = workloads.request_stream(n_req, rate=1.0, seed=4)
    # one-request-per-launch baseline on a sample
    ns = min(64, n_req)
    log(f"serving: one-request-per-launch baseline on {ns} requests")
    t = time.perf_counter()
    for _, rows, single in base_reqs[:ns]:
        s = d4.CudaSearcher(lib, rows=rows, mode=d4.MODE_ENUM if single else d4.MODE_COUNT)
        s.run()
        if single:
            s.solutions(single=True)
        else:
            s.count()
    plain_rps = ns / (time.perf_counter() - t)
    s = None
    out["plain_requests_per_s"] = plain_rps
    for rate in rates:
        reqs = workloads.request_stream(n_req, rate=rate, seed=4)
        lat, wall, res, st = serve_run(d4, lib, reqs, slots)
        tilings = sum(c for c, _ in res.values())
        found = sum(1 for c, sol in res.values() if c)
        row = {"rate": rate, "requests": n_req, "wall_s": wall,
               "requests_per_s": n_req / wall,
               "solutions_per_s": tilings / wall,
               "latency_ms_p50": percentile(lat, 0.5),
               "latency_ms_p99": percentile(lat, 0.99),
               "solvable": found,
               "served_vs_plain": (n_req / wall) / plain_rps,
               "mem_waits": st["mem_waits"]}
        log(f"serve rate {rate}/s: {row['requests_per_s']:.0f} req/s, p50 "
            f"{row['latency_ms_p50']:.1f} ms, p99 {row['latency_ms_p99']:.1f} ms")
        out["rates"].append(row)
    return out


def multi_libraries(d4, insts):
    cache, libs = {}, []
    for shapes, counts, _, w in insts:
        key = (tuple(shapes), tuple(counts), w)
        if key not in cache:
            cache[key] = d4.PieceLibrary(shapes, counts, 0xFF, w)
        libs.append(cache[key])
    return libs


def uneven_section(d4, workloads, n, budget, plain_sample, tune=True):
    """The uneven puzzles (own piece set and width per instance) in one
    multi-library batch with a per-instance state budget, against the same
    instances one per batch; parity against the oracle's counts of the
    instances it finished (tests/golden/uneven_4096_seed3.json)."""
    insts = workloads.uneven_instances(4096, seed=3)[:n]
    t = time.perf_counter()
    libs = multi_libraries(d4, insts)
    t_lib = time.perf_counter() - t
    log(f"uneven: {n} instances, state budget {budget} per instance")
    t = time.perf_counter()
    b = d4.MultiBatch(libs, [x[2] for x in insts])
    t_create = time.perf_counter() - t
    log(f"uneven: {len(set(map(id, libs)))} libraries built on the host in {t_lib:.2f} s, "
        f"batch tables built and uploaded in {t_create:.2f} s")
    b.set_budget(budget)
    if tune:
        t = time.perf_counter()
        tuned = b.tune(64)  # measured launch shape on a sample (outside the timing)
        log(f"uneven autotune {tuned} in {time.perf_counter() - t:.1f} s")
    else:
        tuned = {"threads": 256, "refill": 16}
        b.set_launch(tuned["threads"], tuned["refill"])
    t = time.perf_counter()
    b.run()
    d4.synchronize()
    wall = time.perf_counter() - t
    counts = b.counts()
    peak, fin = b.instances()
    st = b.stats()
    out = {"instances": n, "state_budget": budget, "finished": sum(fin), "wall_s": wall,
           "library_build_s": t_lib, "batch_create_s": t_create,
           "launch": tuned,
           "records_per_s": st["records_extended"] / wall,
           "solutions_per_s": sum(counts) / wall,
           "records_extended": st["records_extended"], "peak_frontier": st["peak_frontier"],
           "splits": st["splits"],
           "extend_GBs": st["extend_bytes"] / (st["extend_ms"] / 1e3) / 1e9
           if st["extend_ms"] else None}
    log(f"uneven batched: {wall:.1f} s, {sum(fin)}/{n} finished, "
        f"{out['records_per_s'] / 1e6:.0f} M records/s")
    fx_path = os.path.join(ROOT, "tests", "golden", "uneven_4096_seed3.json")
    if os.path.exists(fx_path):
        pinned = {int(k): v for k, v in json.load(open(fx_path))["counts"].items()
                  if int(k) < n}
        chk = [i for i in pinned if fin[i]]
        out["parity"] = {"checked_against_oracle": len(chk),
                         "equal": all(counts[i] == pinned[i] for i in chk)}
    if plain_sample:
        m = min(plain_sample, n)
        t = time.perf_counter()
        prec = 0
        for i in range(m):
            pb = d4.MultiBatch([libs[i]], [insts[i][2]])
            pb.set_budget(budget)
            pb.set_launch(tuned["threads"], tuned["refill"])
            pb.run()
            prec += pb.stats()["records_extended"]
        d4.synchronize()
        pdt = time.perf_counter() - t
        out["plain"] = {"instances": m, "records_per_s": prec / pdt, "wall_s": pdt}
        out["batched_vs_plain"] = out["records_per_s"] / (prec / pdt) if prec else None
        log(f"uneven plain: {m} instances one per batch, {prec / pdt / 1e6:.0f} M records/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="run the measured autotuner instead of --launch")
    ap.add_argument("--no-tune", action="store_true",
                    help="library defaults, no --launch; no autotune of the uneven batch")
    ap.add_argument("--launch", default="3,512,16,1536",
                    help="fit_rows,threads,refill,budget of the extend kernel (the "
                         "autotuner's choice for this workload on MI355X)")
    ap.add_argument("--tune-sample", type=int, default=256)
    ap.add_argument("--kernel", choices=("level", "expand", "dfs"), default="level",
                    help="count-mode row-step kernel: level-synchronous (level), partial-row "
                         "snapshots on a per-wave stack (expand) or one depth-first search "
                         "per lane (dfs)")
    ap.add_argument("--tail-rounds", type=int, default=256,
                    help="expand kernel: rounds after a launch's records run out before "
                         "a wave's partial rows move to a follow-up launch")
    ap.add_argument("--serve-rates", default="100,20000",
                    help="comma-separated arrival rates (requests/s); empty to skip")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-slots", type=int, default=1024)
    ap.add_argument("--serve-launch", default="3,256,16,1024",
                    help="fit_rows,threads,refill,budget of the server's extend kernel "
                         "(measured best for the request stream on MI355X)")
    ap.add_argument("--uneven", type=int, default=4096,
                    help="uneven puzzles in the multi-library section (0: skip)")
    ap.add_argument("--uneven-budget", type=int, default=1 << 20)
    ap.add_argument("--uneven-plain", type=int, default=32)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the "
              "instances of this workload shard across GPUs with no exchange, so "
              "N GPUs give N times the one-GPU value", file=sys.stderr)
        sys.exit(2)

    import d4pieces as d4
    import workloads

    if d4.device_count() < 1:
        raise SystemExit("no HIP device visible")
    # device-wide barrier through the library's own HIP runtime (torch ships
    # a separate runtime copy, which is not needed here)
    sync = d4.synchronize

    shapes = workloads.pentominoes()
    boards = workloads.seeded_8x8(args.instances, seed=2)
    lib = d4.PieceLibrary(shapes, [1] * 12, 0xFF, 8)
    tuned = None
    kernel = {"level": d4.KERNEL_LEVEL, "expand": d4.KERNEL_EXPAND,
              "dfs": d4.KERNEL_DFS}[args.kernel]
    lib.set_kernel(kernel, args.tail_rounds)
    if args.launch and not args.tune and not args.no_tune:
        f, t, r, b = (int(x) for x in args.launch.split(","))
        lib.set_launch(f, t, r)
        lib.set_budget(b)
    elif args.tune:
        # measured autotune of the extend launch shape on a sample of the
        # workload (outside the timed region)
        t = time.perf_counter()
        tuned = lib.tune(boards[:args.tune_sample])
        log(f"autotune {tuned} in {time.perf_counter() - t:.1f} s")

    def step():
        b = d4.Batch(lib, boards)
        b.run()
        return b

    for i in range(args.warmup):
        t = time.perf_counter()
        step()
        log(f"warmup step {i} {time.perf_counter() - t:.2f} s")
    sync()
    t0 = time.perf_counter()
    stats, counts = [], None
    for i in range(args.steps):
        t = time.perf_counter()
        b = step()
        stats.append(b.stats())
        log(f"timed step {i} {time.perf_counter() - t:.2f} s")
        c = b.counts()
        if counts is not None and c != counts:
            raise SystemExit("batched counts differ between steps")
        counts = c
    sync()
    dt = time.perf_counter() - t0
    # each engine holds a HIP stream, and streams share the process's few
    # hardware queues: release a section's engines before the next section
    # (the server's two engines only overlap on queues of their own)
    b = None

    rec = sum(s["records_extended"] for s in stats)
    n_sol = sum(counts) * args.steps
    ext_ms = sum(s["extend_ms"] for s in stats)
    ext_launch = sum(s["extend_launches"] for s in stats)
    ext_bytes = sum(s["extend_bytes"] for s in stats)
    achieved = ext_bytes / (ext_ms / 1e3) / 1e9 if ext_ms else 0.0
    traffic, dispatches = None, None
    tpath = os.path.join(ROOT, "profiles", "extend_traffic.json")
    if os.path.exists(tpath):
        prof = json.load(open(tpath))
        traffic = prof.get("hbm_bytes_per_row_step")
        dispatches = prof.get("dispatches_per_row_step")

    out = {
        "metric": METRIC,
        "value": rec / dt,
        "unit": "records extended/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: seeded mt19937_64 boards (seed 2)",
        "config": {
            "workload": f"{args.instances} seeded 8x8 boards, 4 blocked cells, "
                        "12 free pentominoes x1, all tilings, batched",
            "instances": args.instances,
            "board": "8x8",
            "pieces": "12 free pentominoes x1, D4",
            "mode": "count",
            "parallelism": "instances batched in one frontier, 1 GPU",
            "launch": dict(zip(("fit_rows", "threads", "refill"), lib.get_launch()),
                           budget=lib.get_budget(), autotuned=tuned is not None),
        },
        "solutions_per_s": n_sol / dt,
        "solutions_per_step": sum(counts),
        "records_per_step": rec // args.steps,
        "children_per_step": sum(s["children"] for s in stats) // args.steps,
        "peak_frontier": max(s["peak_frontier"] for s in stats),
        "table_reruns_per_step": sum(s["retries"] for s in stats) / args.steps,
        "donated_items_per_step": sum(s["donated"] for s in stats) / args.steps,
        "table_growths_per_step": sum(s["rehashes"] for s in stats) / args.steps,
        "latency_ms": {"batched_p50": dt / args.steps * 1e3,
                       "batched_p99": dt / args.steps * 1e3},
        "roofline": {
            "kernel": "k_extend",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            # one unit = one row step's extend: the main launch plus the
            # load-balance follow-up launches (HIP events around them)
            "per": "row step",
            "avg_launch_ms": ext_ms / max(ext_launch, 1),
            "bytes_per_launch": ext_bytes / max(ext_launch, 1),
            "dispatches_per_row_step": dispatches,
            "compact_GBs": (sum(s["compact_bytes"] for s in stats) /
                            (sum(s["compact_ms"] for s in stats) / 1e3) / 1e9)
            if sum(s["compact_ms"] for s in stats) else None,
        },
    }

    if not args.no_plain:
        n = min(args.plain_sample, len(boards))
        lat, prec, psol = [], 0, 0
        t1 = time.perf_counter()
        for i in range(n):
            s0 = time.perf_counter()
            s = d4.CudaSearcher(lib, rows=boards[i])
            s.run()
            c = s.count()
            sync()
            lat.append((time.perf_counter() - s0) * 1e3)
            if c != counts[i]:
                raise SystemExit(f"plain count differs from batched on board {i}")
            prec += s.stats()["records_extended"]
            psol += c
            log(f"plain board {i} {lat[-1]:.1f} ms")
        pdt = time.perf_counter() - t1
        s = None
        out["plain"] = {
            "boards": n,
            "records_per_s": prec / pdt,
            "solutions_per_s": psol / pdt,
            "latency_ms_p50": percentile(lat, 0.5),
            "latency_ms_p99": percentile(lat, 0.99),
        }
        out["batched_vs_plain"] = (rec / dt) / (prec / pdt) if prec else None
        out["latency_ms"]["plain_p50"] = percentile(lat, 0.5)
        out["latency_ms"]["plain_p99"] = percentile(lat, 0.99)

    if args.serve_rates:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        # the server's frontiers are small (a few requests in flight): it gets
        # a library with its own measured launch shape
        slib = d4.PieceLibrary(shapes, [1] * 12, 0xFF, 8)
        f, t, r, bud = (int(x) for x in args.serve_launch.split(","))
        slib.set_launch(f, t, r)
        slib.set_budget(bud)
        out["served"] = serve_section(d4, slib, workloads, rates, args.serve_requests,
                                      args.serve_slots)
        out["served"]["launch"] = dict(zip(("fit_rows", "threads", "refill", "budget"),
                                           (f, t, r, bud)))
        best = max(out["served"]["rates"], key=lambda r: r["requests_per_s"])
        out["latency_ms"]["served_p50"] = best["latency_ms_p50"]
        out["latency_ms"]["served_p99"] = best["latency_ms_p99"]
        out["served_vs_plain"] = best["served_vs_plain"]

    if args.uneven:
        out["uneven"] = uneven_section(d4, workloads, args.uneven, args.uneven_budget,
                                       args.uneven_plain, tune=not args.no_tune)

    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        base, ok, n_checked = cpu_baseline(boards, counts, shapes, args.cpu_seconds, threads)
        out["cpu_baseline"] = base
        out["parity"] = {"boards_checked_against_oracle": n_checked, "equal": ok}
        if not ok:
            print(json.dumps(out))
            raise SystemExit("GPU counts differ from the oracle")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
