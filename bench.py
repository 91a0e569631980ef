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
  roofline         the dominant kernel's algorithmic HBM bytes / its time
                   (HIP events on the engine's stream) against MI355X's
                   8 TB/s, with the extend, dedup and compaction kernels each
                   under roofline.kernels
  served           Poisson request streams at 8 rates (70% first-solution)
  uneven           4096 puzzles, each with its own piece set and width, in
                   one multi-library batch, then the ones a state budget
                   dropped in a second pass with a larger budget
  large            12 pentominoes x2 on 8x15 (one bounded-memory search),
                   and a sample of its frontier checked state by state
                   against the oracle
  cpu_baseline     the oracle (CPU restatement of the reference backtracker)
                   on the host's cores: the 256 boards, 6x10 single-threaded,
                   the request stream offline, and the large search's rate
                   on its sampled states

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import statistics
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


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle
    return Oracle(os.path.join(ROOT, "oracle", "liboracle.so"))


# --------------------------------------------------------------------------
# CPU baseline (the oracle, test infrastructure only: it is the checker and
# the reported baseline, never part of the GPU measurement)
# --------------------------------------------------------------------------

def cpu_boards(oracle, boards, gpu_counts, shapes, threads, seconds):
    """solve_count on the boards, one board per task on `threads` threads,
    until all are done or `seconds` pass."""
    done, sols, ok = 0, 0, True
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while done < len(boards) and time.perf_counter() - t0 < seconds:
            chunk = boards[done:done + 4 * threads]
            res = list(ex.map(lambda b: oracle.count(shapes, [1] * 12, b), chunk))
            for j, c in enumerate(res):
                ok &= c == gpu_counts[done + j]
            sols += sum(res)
            done += len(chunk)
            log(f"cpu baseline: {done}/{len(boards)} boards {time.perf_counter() - t0:.1f} s")
    dt = time.perf_counter() - t0
    return {"boards": done, "of": len(boards), "wall_s": dt, "solutions_per_s": sols / dt,
            "boards_per_s": done / dt, "threads": threads, "parity_equal": ok}


def cpu_plain_6x10(oracle, workloads):
    shapes, counts, rows, _ = workloads.plain_6x10()
    t = time.perf_counter()
    n = oracle.count(shapes, counts, rows)
    t_count = time.perf_counter() - t
    t = time.perf_counter()
    sols = oracle.solve(shapes, counts, rows)
    t_set = time.perf_counter() - t
    return {"count": n, "solutions": len(sols), "count_s": t_count, "full_set_s": t_set,
            "threads": 1}


def cpu_requests(oracle, workloads, n_req, threads):
    """The request stream offline: every request solved by the oracle (first
    solution or count), one request per task."""
    shapes = workloads.pentominoes()
    reqs = workloads.request_stream(n_req, rate=1.0, seed=4)

    def one(r):
        _, rows, single = r
        if single:
            return len(oracle.solve(shapes, [1] * 12, rows, single=True))
        return oracle.count(shapes, [1] * 12, rows)

    t = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(one, reqs))
    dt = time.perf_counter() - t
    return {"requests": n_req, "wall_s": dt, "requests_per_s": n_req / dt, "threads": threads}


# --------------------------------------------------------------------------
# serving
# --------------------------------------------------------------------------

def serve_run(d4, lib, reqs, slots):
    """Open-loop replay of a request stream (arrival times in seconds from
    the start) through one Server; returns per-request latencies (ms), wall
    time, results, server stats and the host time spent in server steps."""
    srv = d4.Server(lib, slots=slots, max_height=8)
    n = len(reqs)
    t_arr, sub, res, lat, last_log = {}, 0, {}, {}, -1
    rid_of = {}
    step_s, steps = 0.0, 0
    t0 = time.perf_counter()
    while len(res) < n:
        now = time.perf_counter() - t0
        while sub < n and reqs[sub][0] <= now:
            rid = srv.submit(reqs[sub][1], reqs[sub][2])
            t_arr[rid] = reqs[sub][0]
            rid_of[rid] = sub
            sub += 1
        if sub > len(res):
            ts = time.perf_counter()
            srv.step()
            step_s += time.perf_counter() - ts
            steps += 1
        while True:
            r = srv.poll()
            if r is None:
                break
            res[r[0]] = (r[1], r[2])
            lat[rid_of[r[0]]] = (time.perf_counter() - t0 - t_arr[r[0]]) * 1e3
        if len(res) and len(res) % 250 == 0 and len(res) != last_log:
            last_log = len(res)
            log(f"  served {len(res)}/{n}")
        if sub == len(res) and sub < n:
            # idle until the next arrival
            time.sleep(max(0.0, reqs[sub][0] - (time.perf_counter() - t0)))
    return lat, time.perf_counter() - t0, res, srv.stats(), step_s, steps


def serve_section(d4, lib, workloads, rates, n_req, slots):
    """Served requests (Poisson arrivals, 70% first-solution) against the
    same requests solved one per CudaSearcher, one after another."""
    out = {"rates": []}
    base_reqs = workloads.request_stream(n_req, rate=1.0, seed=4)
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
    bands = ((30, 39), (40, 49), (50, 64))
    # untimed warmup, the whole stream at the highest rate (all its requests
    # in one burst: the largest batches): the server's buffers grow to their
    # working sizes once, and later servers take them from the allocator's
    # cache instead of allocating inside the measured stream
    if rates:
        warm = workloads.request_stream(n_req, rate=max(rates), seed=6)
        t = time.perf_counter()
        serve_run(d4, lib, warm, slots)
        log(f"serving: warmup of {len(warm)} requests in {time.perf_counter() - t:.2f} s")
    for rate in rates:
        reqs = workloads.request_stream(n_req, rate=rate, seed=4)
        lat, wall, res, st, step_s, steps = serve_run(d4, lib, reqs, slots)
        tilings = sum(c for c, _ in res.values())
        found = sum(1 for c, sol in res.values() if c)
        by_kind = {}
        for kind, want in (("first_solution", True), ("all_solutions", False)):
            v = [lat[i] for i in range(n_req) if reqs[i][2] == want]
            by_kind[kind] = {"requests": len(v), "p50_ms": percentile(v, 0.5),
                             "p99_ms": percentile(v, 0.99)}
        by_band = {}
        for lo, hi in bands:
            v = [lat[i] for i in range(n_req)
                 if lo <= sum(bin(r).count("1") for r in reqs[i][1]) <= hi]
            by_band[f"{lo}-{hi}_free_cells"] = {"requests": len(v), "p50_ms": percentile(v, 0.5),
                                                "p99_ms": percentile(v, 0.99)}
        lat_all = [lat[i] for i in range(n_req)]
        gpu_ms = st["all"]["gpu_ms"] + st["one"]["gpu_ms"]
        row = {"rate": rate, "requests": n_req, "wall_s": wall,
               "requests_per_s": n_req / wall,
               "keeps_up": n_req / wall >= 0.95 * rate,
               "solutions_per_s": tilings / wall,
               "latency_ms_p50": percentile(lat_all, 0.5),
               "latency_ms_p99": percentile(lat_all, 0.99),
               "by_kind": by_kind, "by_free_cells": by_band,
               "server_steps": steps,
               "host_ms_per_step": step_s / max(steps, 1) * 1e3,
               # the two engines' device time (they overlap on the GPU)
               "gpu_ms_per_step": gpu_ms / max(steps, 1),
               "solvable": found,
               "served_vs_plain": (n_req / wall) / plain_rps,
               "mem_waits": st["mem_waits"]}
        log(f"serve rate {rate:g}/s: {row['requests_per_s']:.0f} req/s, p50 "
            f"{row['latency_ms_p50']:.1f} ms, p99 {row['latency_ms_p99']:.1f} ms, "
            f"host {row['host_ms_per_step']:.2f} / gpu {row['gpu_ms_per_step']:.2f} ms per step")
        out["rates"].append(row)
    return out


# --------------------------------------------------------------------------
# uneven instances (multi-library batch)
# --------------------------------------------------------------------------

def multi_libraries(d4, insts):
    cache, libs = {}, []
    for shapes, counts, _, w in insts:
        key = (tuple(shapes), tuple(counts), w)
        if key not in cache:
            cache[key] = d4.PieceLibrary(shapes, counts, 0xFF, w)
        libs.append(cache[key])
    return libs


def uneven_section(d4, workloads, n, budget, budget2, plain_sample, tune=True,
                   pass2_seconds=60.0, kernel="level"):
    """The uneven puzzles (own piece set and width per instance) in one
    multi-library batch with a per-instance state budget; the instances that
    budget drops run again in a second batch with budget2; the rest is
    reported with the frontier size at which it stopped.  Against the same
    instances one per batch, and the oracle's counts of the instances it
    finished (tests/golden/uneven_4096_seed3.json)."""
    insts = workloads.uneven_instances(4096, seed=3)[:n]
    t = time.perf_counter()
    libs = multi_libraries(d4, insts)
    t_lib = time.perf_counter() - t
    log(f"uneven: {n} instances, state budget {budget} per instance")
    # the per-library tables built on the host and uploaded, against built on
    # the device from the uploaded candidates (the batch that runs)
    t = time.perf_counter()
    bh = d4.MultiBatch(libs, [x[2] for x in insts], host_tables=True)
    t_create_host = time.perf_counter() - t
    bh = None
    t = time.perf_counter()
    kid = d4.KERNEL_LEVEL if kernel == "level" else d4.KERNEL_DFS
    b = d4.MultiBatch(libs, [x[2] for x in insts], kernel=kid)
    t_create = time.perf_counter() - t
    log(f"uneven: {len(set(map(id, libs)))} libraries built on the host in {t_lib:.2f} s, "
        f"batch created with device-built tables in {t_create:.3f} s "
        f"(host-built tables: {t_create_host:.3f} s)")
    b.set_budget(budget)
    # (the launch shape is the depth-first kernel's; the level kernel's
    # launches are sized from the frontier)
    tuned = None
    if kernel == "dfs" and tune:
        t = time.perf_counter()
        tuned = b.tune(64)  # measured launch shape on a sample (outside the timing)
        log(f"uneven autotune {tuned} in {time.perf_counter() - t:.1f} s")
    elif kernel == "dfs":
        tuned = {"threads": 256, "refill": 16}
        b.set_launch(tuned["threads"], tuned["refill"])

    def shape(x):
        if tuned:
            x.set_launch(tuned["threads"], tuned["refill"])
    # an untimed cold pass of the same batch first (reported beside): it
    # allocates the device blocks of the frontiers, tables and level buffers,
    # which the timed pass takes from the library's cache (device memory just
    # released to the driver, as by the cache release before this section,
    # becomes available to a new allocation only seconds later)
    bw = d4.MultiBatch(libs, [x[2] for x in insts], kernel=kid)
    bw.set_budget(budget)
    shape(bw)
    t = time.perf_counter()
    bw.run()
    d4.synchronize()
    cold = time.perf_counter() - t
    bw = None
    log(f"uneven cold pass (allocating): {cold:.1f} s")
    t = time.perf_counter()
    b.run()
    d4.synchronize()
    wall = time.perf_counter() - t
    counts = b.counts()
    peak, fin = b.instances()
    st = b.stats()
    b = None
    out = {"instances": n, "state_budget": budget, "finished": sum(fin), "wall_s": wall,
           "cold_wall_s": cold,
           "library_build_s": t_lib, "batch_create_s": t_create,
           "batch_create_host_tables_s": t_create_host,
           "launch": {"kernel": kernel, **(tuned or {})},
           "records_per_s": st["records_extended"] / wall,
           "solutions_per_s_finished": sum(counts) / wall,
           "records_extended": st["records_extended"], "peak_frontier": st["peak_frontier"],
           "splits": st["splits"],
           "extend_GBs": st["extend_bytes"] / (st["extend_ms"] / 1e3) / 1e9
           if st["extend_ms"] else None}
    log(f"uneven batched: {wall:.1f} s, {sum(fin)}/{n} finished, "
        f"{out['records_per_s'] / 1e6:.0f} M records/s")
    # second pass: the dropped instances with a larger budget, in batches of
    # 128 (one progress line each) until pass2_seconds have gone by
    rest = [i for i in range(n) if not fin[i]]
    if rest and budget2 > budget:
        t = time.perf_counter()
        ran, fin2_n, rec2, splits2 = 0, 0, 0, 0
        for c0 in range(0, len(rest), 128):
            if time.perf_counter() - t > pass2_seconds:
                break
            part = rest[c0:c0 + 128]
            b2 = d4.MultiBatch([libs[i] for i in part], [insts[i][2] for i in part],
                               kernel=kid)
            b2.set_budget(budget2)
            shape(b2)
            b2.run()
            c2 = b2.counts()
            peak2, fin2 = b2.instances()
            st2 = b2.stats()
            b2 = None
            for j, i in enumerate(part):
                if fin2[j]:
                    counts[i], fin[i] = c2[j], True
                peak[i] = max(peak[i], peak2[j])
            ran += len(part)
            fin2_n += sum(fin2)
            rec2 += st2["records_extended"]
            splits2 += st2["splits"]
            log(f"uneven second pass: {ran}/{len(rest)} instances, {fin2_n} finished, "
                f"{time.perf_counter() - t:.1f} s")
        d4.synchronize()
        wall2 = time.perf_counter() - t
        out["second_pass"] = {"instances": len(rest), "ran": ran, "state_budget": budget2,
                              "finished": fin2_n, "wall_s": wall2, "records_extended": rec2,
                              "splits": splits2, "batch": 128}
    out["finished_total"] = sum(fin)
    # (states_at_stop: the largest frontier share the instance reached in the
    # pass that dropped it; an instance the time bound left unrun keeps the
    # first pass's figure)
    out["unfinished"] = [{"instance": i, "states_at_stop": peak[i]}
                         for i in range(n) if not fin[i]]
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
            pb = d4.MultiBatch([libs[i]], [insts[i][2]], kernel=kid)
            pb.set_budget(budget)
            shape(pb)
            pb.run()
            prec += pb.stats()["records_extended"]
        d4.synchronize()
        pdt = time.perf_counter() - t
        out["plain"] = {"instances": m, "records_per_s": prec / pdt, "wall_s": pdt}
        out["batched_vs_plain"] = out["records_per_s"] / (prec / pdt) if prec else None
        log(f"uneven plain: {m} instances one per batch, {prec / pdt / 1e6:.0f} M records/s")
    return out


# --------------------------------------------------------------------------
# large single enumeration
# --------------------------------------------------------------------------

def sub_board(rows, g, key, maxh, layout, counts):
    """The board and piece counts left to a frontier state of step g."""
    occ_bits = 8 * (maxh - 1)
    occ = key & ((1 << occ_bits) - 1)
    sub = list(rows[g:])
    for j in range(min(maxh - 1, len(sub))):
        sub[j] &= ~(occ >> (8 * j)) & 0xFF
    left = []
    for (off, w), c in zip(layout, counts):
        used = (key >> (occ_bits + off)) & ((1 << w) - 1) if w else 0
        left.append(c - used)
    return sub, left


def step_modes_section(d4, workloads, boards, expected, serve_slots, n_req=300, rate=200.0):
    """The same work with count-mode row steps issued three ways: host round
    trips between the parts of a step, one queued chain per step, and steps
    replayed from recorded HIP graphs (many per round trip).  Per mode: the
    batched 256-board solve, the 6x10 board alone (latency), and a served
    stream (the count-mode requests' engine steps that way)."""
    shapes = workloads.pentominoes()
    s6, c6, r6, _ = workloads.plain_6x10()
    reqs = workloads.request_stream(n_req, rate=rate, seed=5)
    # (the serving sweep before this section leaves the allocator's cache
    # full of its servers' blocks: released here, so that the untimed first
    # solve of each mode allocates what the timed ones reuse, instead of a
    # timed one releasing the cache and waiting for the freed memory)
    d4.release_cached()
    out = {}
    for name, mode in (("general", d4.STEP_GENERAL), ("fused", d4.STEP_FUSED),
                       ("graph", d4.STEP_GRAPH)):
        lib = d4.PieceLibrary(shapes, [1] * 12, 0xFF, 8)
        lib.set_step_mode(mode)
        ts, st = [], None
        for it in range(4):
            b = d4.Batch(lib, boards)
            t = time.perf_counter()
            b.run()
            dt = time.perf_counter() - t
            if b.counts() != expected:
                raise SystemExit(f"step mode {name}: batched counts differ")
            if it:
                ts.append(dt * 1e3)
                st = b.stats()
            b = None
        lib6 = d4.PieceLibrary(s6, c6, 0xFF, 6)
        lib6.set_step_mode(mode)
        t6, st6 = [], None
        for it in range(23):
            sr = d4.CudaSearcher(lib6, rows=r6)
            t = time.perf_counter()
            sr.run()
            n6 = sr.count()
            dt = time.perf_counter() - t
            if n6 != 9356:
                raise SystemExit(f"step mode {name}: 6x10 count {n6}")
            if it >= 3:
                t6.append(dt * 1e3)
                st6 = sr.stats()
            sr = None
        slib = d4.PieceLibrary(shapes, [1] * 12, 0xFF, 8)
        slib.set_step_mode(mode)
        lat, wall, res, sst, step_s, steps = serve_run(d4, slib, reqs, serve_slots)
        rows = st["rows_done"]
        out[name] = {
            "batched_ms": statistics.median(ts),
            "batched_host_syncs_per_row_step": st["host_syncs"] / rows,
            "batched_gpu_ms": st["gpu_ms"],
            "plain_6x10_p50_ms": percentile(t6, 0.5),
            "plain_6x10_p99_ms": percentile(t6, 0.99),
            "plain_6x10_gpu_ms": st6["gpu_ms"],
            "plain_6x10_host_syncs": st6["host_syncs"],
            "plain_6x10_graph_replays": st6["graph_replays"],
            "served_rate": rate,
            "served_p50_ms": percentile(list(lat.values()), 0.5),
            "served_p99_ms": percentile(list(lat.values()), 0.99),
            "served_requests_per_s": len(res) / wall,
            "served_count_engine_host_syncs_per_step":
                sst["all"]["host_syncs"] / max(1, sst["all"]["rows_done"]),
        }
        log(f"step mode {name}: batched {out[name]['batched_ms']:.1f} ms "
            f"({out[name]['batched_host_syncs_per_row_step']:.2f} syncs/row step), 6x10 p50 "
            f"{out[name]['plain_6x10_p50_ms']:.2f} ms, served p50 "
            f"{out[name]['served_p50_ms']:.1f} / p99 {out[name]['served_p99_ms']:.1f} ms")
    return out


def large_section(d4, workloads, oracle, sample_height, n_sample, cpu_seconds):
    """12 pentominoes x2 on 8x15: one count-mode search under the default
    memory cap (80% of free device memory; a row that would not fit parks
    half of its frontier), timed; a sample of its frontier at sample_height
    rows left is counted state by state on the GPU and by the oracle."""
    shapes, counts, rows, w = workloads.large_8x15()
    lib = d4.PieceLibrary(shapes, counts, 0xFF, w)
    # one line per row on stderr from the engine (progress of a long search)
    os.environ["D4_TRACE"] = "1"
    s = d4.CudaSearcher(lib, rows=rows)
    del os.environ["D4_TRACE"]
    s.set_sample(sample_height, n_sample)
    d4.synchronize()
    t = time.perf_counter()
    s.run()
    d4.synchronize()
    wall = time.perf_counter() - t
    total = s.count()
    st = s.stats()
    sample = s.sample()
    s = None
    out = {"workload": "12 free pentominoes x2 (24 pieces, 120 cells) on 8 wide x 15 rows",
           "count": total, "wall_s": wall, "records_extended": st["records_extended"],
           "records_per_s": st["records_extended"] / wall, "solutions_per_s": total / wall,
           "peak_frontier": st["peak_frontier"], "splits": st["splits"],
           "rows": st["rows_done"], "table_growths": st["rehashes"]}
    log(f"large 8x15: {total} tilings in {wall:.1f} s, peak frontier {st['peak_frontier']}, "
        f"{st['splits']} splits")
    # state by state: the GPU and the oracle on what each sampled state leaves
    maxh = lib.max_height
    layout = lib.counter_layout()
    g = len(rows) - sample_height
    checked, equal, cpu_s, cpu_tilings, gpu_s = 0, True, 0.0, 0, 0.0
    for key, cnt in sample:
        if cpu_s > cpu_seconds:
            break
        sub, left = sub_board(rows, g, key, maxh, layout, counts)
        t = time.perf_counter()
        want = oracle.count(shapes, left, sub)
        cpu_s += time.perf_counter() - t
        t = time.perf_counter()
        sl = d4.PieceLibrary(shapes, left, 0xFF, w)
        ss = d4.CudaSearcher(sl, rows=sub)
        ss.run()
        got = ss.count()
        gpu_s += time.perf_counter() - t
        equal &= got == want
        checked += 1
        cpu_tilings += cnt * want
    out["sample"] = {"height": sample_height, "states": checked, "gpu_equals_oracle": equal,
                     "oracle_s": cpu_s, "tilings_through_sample": cpu_tilings}
    log(f"large sample: {checked} states at height {sample_height}, GPU == oracle: {equal}, "
        f"oracle {cpu_s:.1f} s")
    return out, (cpu_tilings / cpu_s if cpu_s else None)


# --------------------------------------------------------------------------
# main
# --------------------------------------------------------------------------

def kernel_roofline(name, bytes_, ms, traffic=None):
    gbs = bytes_ / (ms / 1e3) / 1e9 if ms else 0.0
    return {"kernel": name, "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
            "ms_per_row_step": None, "bytes_per_row_step": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="wall-time bound of the oracle on the 256 boards")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every host core")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="run the measured autotuner instead of --launch")
    ap.add_argument("--no-tune", action="store_true",
                    help="library defaults, no --launch; no autotune of the uneven batch")
    ap.add_argument("--launch", default="3,512,16,1536",
                    help="fit_rows,threads,refill,budget of the depth-first kernel (the "
                         "autotuner's choice for this workload on MI355X; the level "
                         "kernel uses the fit rows)")
    ap.add_argument("--tune-sample", type=int, default=256)
    ap.add_argument("--kernel", choices=("level", "expand", "dfs"), default="level",
                    help="count-mode row-step kernel: level-synchronous (level), partial-row "
                         "snapshots on a per-wave stack (expand) or one depth-first search "
                         "per lane (dfs)")
    ap.add_argument("--mbatch-kernel", choices=("level", "dfs"), default="level",
                    help="row-step kernel of the multi-library batches: level-synchronous "
                         "(level) or one depth-first search per lane (dfs)")
    ap.add_argument("--step-mode", choices=("general", "fused", "graph"), default="fused",
                    help="how count-mode level row steps are issued: host round trips "
                         "between their parts, one queued chain, or that chain replayed "
                         "from a HIP graph")
    ap.add_argument("--no-step-modes", action="store_true",
                    help="skip the general / fused / graph-replay comparison")
    ap.add_argument("--tail-rounds", type=int, default=256,
                    help="expand kernel: rounds after a launch's records run out before "
                         "a wave's partial rows move to a follow-up launch")
    ap.add_argument("--serve-rates", default="100,200,500,1000,2000,5000,10000,20000",
                    help="comma-separated arrival rates (requests/s); empty to skip")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-slots", type=int, default=1024)
    ap.add_argument("--serve-launch", default="3,256,16,1024",
                    help="fit_rows,threads,refill,budget of the server's extend kernel "
                         "(measured best for the request stream on MI355X)")
    ap.add_argument("--uneven", type=int, default=4096,
                    help="uneven puzzles in the multi-library section (0: skip)")
    ap.add_argument("--uneven-budget", type=int, default=1 << 20)
    ap.add_argument("--uneven-budget2", type=int, default=1 << 22,
                    help="state budget of the second pass over the dropped instances")
    ap.add_argument("--uneven-plain", type=int, default=32)
    ap.add_argument("--large", type=int, default=1, help="run the 8x15 search (0: skip)")
    ap.add_argument("--large-sample-height", type=int, default=4)
    ap.add_argument("--large-sample", type=int, default=512)
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
    # device-wide barrier through the library's own HIP runtime
    sync = d4.synchronize

    shapes = workloads.pentominoes()
    boards = workloads.seeded_8x8(args.instances, seed=2)
    lib = d4.PieceLibrary(shapes, [1] * 12, 0xFF, 8)
    tuned = None
    kernel = {"level": d4.KERNEL_LEVEL, "expand": d4.KERNEL_EXPAND,
              "dfs": d4.KERNEL_DFS}[args.kernel]
    lib.set_kernel(kernel, args.tail_rounds)
    step_mode = {"general": d4.STEP_GENERAL, "fused": d4.STEP_FUSED,
                 "graph": d4.STEP_GRAPH}[args.step_mode]
    lib.set_step_mode(step_mode)
    if args.launch and not args.tune and not args.no_tune:
        f, t, r, b = (int(x) for x in args.launch.split(","))
        lib.set_launch(f, t, r)
        lib.set_budget(b)
    elif args.tune:
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
        log(f"warmup step {i} {time.perf_counter() - t:.3f} s")
    sync()
    t0 = time.perf_counter()
    stats, counts = [], None
    for i in range(args.steps):
        t = time.perf_counter()
        b = step()
        stats.append(b.stats())
        c = b.counts()
        # a batch holds its engine's buffers: drop it before the next step
        # so that step reuses them from the allocator's cache
        b = None
        log(f"timed step {i} {time.perf_counter() - t:.3f} s")
        if counts is not None and c != counts:
            raise SystemExit("batched counts differ between steps")
        counts = c
    sync()
    dt = time.perf_counter() - t0
    # each engine holds a HIP stream, and streams share the process's few
    # hardware queues: release a section's engines before the next section
    b = None

    def tot(k):
        return sum(s[k] for s in stats)

    rec = tot("records_extended")
    n_sol = sum(counts) * args.steps
    rows = tot("rows_done")
    prof = {}
    tpath = os.path.join(ROOT, "profiles", "row_step_traffic.json")
    if os.path.exists(tpath):
        prof = json.load(open(tpath))
    kernels = {}
    for name, bkey, mkey in (("extend", "extend_bytes", "extend_ms"),
                             ("dedup", "dedup_bytes", "dedup_ms"),
                             ("compact", "compact_bytes", "compact_ms")):
        if tot(mkey):
            k = kernel_roofline(name, tot(bkey), tot(mkey),
                                prof.get(name, {}).get("hbm_bytes_per_row_step"))
            k["ms_per_row_step"] = tot(mkey) / rows
            k["bytes_per_row_step"] = tot(bkey) / rows
            kernels[name] = k
    if "dedup" in kernels:
        # Row steps with many children dedup through hash buckets merged in
        # LDS (a histogram, a scan and two scatter passes that stream, then
        # one workgroup per bucket); small ones through the device-memory
        # table, whose random 64-bit atomics the device runs at a fixed rate
        # whatever the table's size -- measured here, it is the ceiling that
        # made the buckets worth their extra passes.
        rate = d4.atomic_rate()
        kernels["dedup"]["row_steps_bucketed"] = tot("part_steps")
        kernels["dedup"]["row_steps_table"] = rows - tot("part_steps")
        kernels["dedup"]["children_per_s"] = tot("children") / (tot("dedup_ms") / 1e3)
        kernels["dedup"]["random_atomics_per_s_measured"] = rate
        log(f"dedup: {tot('part_steps')} of {rows} row steps bucketed, "
            f"{kernels['dedup']['children_per_s'] / 1e9:.1f} G kept children/s; the device runs "
            f"{rate / 1e9:.1f} G random atomics/s")
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_row_step"])
    roof = dict(kernels[dominant])
    roof["per"] = "row step"
    roof["kernels"] = kernels

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
            "kernel": args.kernel,
            "step_mode": args.step_mode,
            "parallelism": "instances batched in one frontier, 1 GPU",
            "launch": dict(zip(("fit_rows", "threads", "refill"), lib.get_launch()),
                           budget=lib.get_budget(), autotuned=tuned is not None),
        },
        "solutions_per_s": n_sol / dt,
        "solutions_per_step": sum(counts),
        "records_per_step": rec // args.steps,
        "children_per_step": tot("children") // args.steps,
        "peak_frontier": max(s["peak_frontier"] for s in stats),
        "table_reruns_per_step": tot("retries") / args.steps,
        "chunk_reruns_per_step": tot("chunk_reruns") / args.steps,
        "host_syncs_per_row_step": tot("host_syncs") / rows,
        "fused_steps_per_step": tot("fused_steps") / args.steps,
        "fused_fallbacks_per_step": tot("fused_fallbacks") / args.steps,
        "table_growths_per_step": tot("rehashes") / args.steps,
        "bucketed_dedup_steps_per_step": tot("part_steps") / args.steps,
        "latency_ms": {"batched_p50": dt / args.steps * 1e3,
                       "batched_p99": dt / args.steps * 1e3},
        "roofline": roof,
    }

    # the large search runs while the device's memory is still unfragmented
    oracle = None if args.no_cpu else load_oracle()
    large_rate = None
    if args.large:
        if oracle is None:
            oracle = load_oracle()
        out["large"], large_rate = large_section(d4, workloads, oracle, args.large_sample_height,
                                                 args.large_sample, 30.0)
        # (the next sections' small allocations would otherwise release the
        # hundreds of GB the allocator caches from this search all at once,
        # inside a timed region)
        d4.release_cached()

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
        pdt = time.perf_counter() - t1
        s = None
        log(f"plain: {n} boards one per CudaSearcher, p50 {percentile(lat, 0.5):.1f} ms")
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
        slib.set_kernel(kernel, args.tail_rounds)
        slib.set_step_mode(step_mode)
        f, t, r, bud = (int(x) for x in args.serve_launch.split(","))
        slib.set_launch(f, t, r)
        slib.set_budget(bud)
        out["served"] = serve_section(d4, slib, workloads, rates, args.serve_requests,
                                      args.serve_slots)
        out["served"]["launch"] = dict(zip(("fit_rows", "threads", "refill", "budget"),
                                           (f, t, r, bud)))
        rows_ = out["served"]["rates"]
        keep = [x for x in rows_ if x["keeps_up"]] or rows_[:1]
        head = max(keep, key=lambda x: x["rate"])  # highest rate served as it arrives
        sat = max(rows_, key=lambda x: x["requests_per_s"])
        out["served"]["headline_rate"] = head["rate"]
        out["latency_ms"]["served_rate"] = head["rate"]
        out["latency_ms"]["served_p50"] = head["latency_ms_p50"]
        out["latency_ms"]["served_p99"] = head["latency_ms_p99"]
        out["served_vs_plain"] = sat["served_vs_plain"]
        out["served_saturated_requests_per_s"] = sat["requests_per_s"]

    if not args.no_step_modes and kernel == d4.KERNEL_LEVEL:
        out["step_modes"] = step_modes_section(d4, workloads, boards, counts, args.serve_slots)

    if args.uneven:
        # (a clean device for the multi-library batch: it splits its frontier
        # when an allocation fails, and blocks cached by the sections before
        # it, sized for their own buffers, can leave too little room)
        freed = d4.release_cached()
        log(f"uneven: {freed / 2**30:.1f} GiB of cached device memory released first")
        out["uneven"] = uneven_section(d4, workloads, args.uneven, args.uneven_budget,
                                       args.uneven_budget2, args.uneven_plain,
                                       tune=not args.no_tune, kernel=args.mbatch_kernel)


    if not args.no_cpu:
        threads = args.cpu_threads or host_cores()
        c2 = cpu_boards(oracle, boards, counts, shapes, threads, args.cpu_seconds)
        per = {"seeded_8x8_256": c2, "plain_6x10": cpu_plain_6x10(oracle, workloads)}
        log(f"cpu baseline: 6x10 count {per['plain_6x10']['count_s']:.2f} s, "
            f"full set {per['plain_6x10']['full_set_s']:.2f} s (1 thread)")
        if args.serve_rates:
            per["requests_offline"] = cpu_requests(oracle, workloads, args.serve_requests,
                                                   threads)
            log(f"cpu baseline: {args.serve_requests} requests offline in "
                f"{per['requests_offline']['wall_s']:.1f} s")
        if large_rate:
            tot_l = out["large"]["count"]
            per["large_8x15"] = {
                "method": "oracle on the sampled frontier states' remaining boards, one "
                          "thread; rate = tilings through the sample (state count x its "
                          "completions) / oracle time",
                "tilings_per_s": large_rate,
                "estimated_single_thread_s": tot_l / large_rate,
                "states": out["large"]["sample"]["states"], "threads": 1}
        out["cpu_baseline"] = {
            "value": c2["solutions_per_s"],
            "unit": "solutions/s",
            "cores": threads,
            "cpu_model": cpu_model(),
            "kind": "port",
            "sample": f"{c2['boards']} of {len(boards)} boards of the workload, one board per "
                      f"task on {threads} threads, {c2['wall_s']:.1f} s wall",
            "per_config": per,
        }
        out["vs_cpu"] = out["solutions_per_s"] / c2["solutions_per_s"]
        out["parity"] = {"boards_checked_against_oracle": c2["boards"],
                         "equal": c2["parity_equal"]}
        if not c2["parity_equal"]:
            print(json.dumps(out))
            raise SystemExit("GPU counts differ from the oracle")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
