#!/usr/bin/env python3
"""Benchmark of the GPU GROUP BY path on one MI355X.

With no --workload, one run measures every configuration and prints ONE
JSON line: the headline (the filtered query below) at the top level, and
`configs` holding a record per other configuration - plain (4M rows),
highcard (200M rows, ~149M groups), wide (100M rows grouped by two
full-range BIGINT keys, ~43M groups), batched (512 uneven queries in one
launch) and serving (Poisson arrivals) - each with its own value, roofline
and CPU baseline, plus `measured_peak`: the device's read and copy bandwidth
measured in the same run (the roofline's second denominator).

Headline workload (the filtered scan -> filter -> GROUP BY -> top-N query on
200M rows in 1M-row segments / 65,536-row row groups):
  SELECT region, COUNT(*), SUM(amount), AVG(price) FROM t
  WHERE ts >= T0+10d AND ts < T0+17d
  GROUP BY region ORDER BY SUM(amount) DESC LIMIT 100
region INT Zipf(1.1) over 10,000 ids, ts DATETIME over 30 days, amount
DECIMAL(18,2) log-normal cents, price DOUBLE U[0,1000).

A step is one execution of the query over the HBM-resident table (columns
staged once, before timing).  `value` is rows scanned per second.  The
roofline block prices the dominant kernel (the fused scan/filter/aggregate)
at 28 algorithmic bytes per row (4+8+8+8) against 8 TB/s, with its duration
from HIP events on the plan's stream.  For `highcard` it prices the radix
partition/aggregate kernels together (events around the sequence) at 21
bytes per row plus 33 bytes of group state per group; for `wide` the hash
partition kernels at 20 bytes per row plus 32 per group (128-bit key, COUNT,
SUM).  CPU baselines time
the oracle (the C restatement of thinDB's grouped aggregate) on a bounded
sample of the same table, on every host core this job has (at most 16, the
box's share per GPU) and on one core.

Workload `batched`: 512 GROUP BY queries over 512 tables of uneven size
(rows log-uniform in [1e4, 2e7], key cardinality log-uniform in [4, 1e6],
half of them with a 50%-selective range filter), all run as one batch (one
persistent launch, replayed as a captured HIP graph, strategies autotuned per
(rows, groups) bucket) against the same 512 plans executed one after another
through the single-query operator.  A step is one pass over all 512 queries,
results delivered to host memory; `value` is queries per second.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3]
       [--workload all|filtered|plain|highcard|wide|batched|serving] [--having]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0

WORKLOADS = {
    # name: rows, algorithmic bytes per row, description
    "filtered": (200_000_000, 28, "filtered GROUP BY region top-100 by SUM(amount), 200M rows"),
    "plain": (4_000_000, 16, "GROUP BY k COUNT/SUM/MIN/MAX, 4M rows"),
    "highcard": (200_000_000, 21, "GROUP BY (user_id, bucket) COUNT/SUM/MAX top-10 by COUNT, 200M rows"),
    "wide": (100_000_000, 20, "GROUP BY (a, b) over two full-range BIGINT keys COUNT/SUM top-10 by COUNT, 100M rows"),
    "batched": (0, 12, "512 GROUP BY key COUNT/SUM/MIN/MAX queries over 512 uneven tables, one batch"),
    "serving": (0, 0, "serving loop under Poisson arrivals"),
}
BATCH_SEED = 7
HIGHCARD_GROUP_BYTES = 33  # key 9 + COUNT 8 + SUM 8 + MAX 8 per group
WIDE_GROUP_BYTES = 32      # key 16 + COUNT 8 + SUM 8 per group


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# --having: the HAVING each workload adds (groups filtered before ORDER BY / LIMIT)
HAVING = {
    "filtered": "HAVING AVG(price) > 500",
    "plain": "HAVING COUNT(*) > 3900 AND MAX(v) > 0",
    "highcard": "HAVING COUNT(*) >= 2 AND SUM(v) > 0",
}


# --distinct: GROUP BY bucket, COUNT(DISTINCT user_id) over the high-cardinality table
DISTINCT_DESC = "GROUP BY bucket COUNT(*), COUNT(DISTINCT {col}) top-10 by the distinct count, 200M rows"
DISTINCT_BYTES_PER_ROW = 9  # bucket 1 + user_id 8


def make_request(workload: str, having: bool = False, distinct: bool = False, where_tree: bool = False,
                 distinct_col: str = "user_id"):
    from thindb_amd import gen
    from thindb_amd.exec import And, AggSpec, Not, Or, Predicate, Request, SortSpec
    if distinct:
        return Request(["bucket"], [AggSpec("count", None, "c"), AggSpec("count_distinct", distinct_col, "u")],
                       [SortSpec("u", True)], 10)
    if workload == "filtered":
        where = And((Predicate("ts", "gte", gen.FILTER_LO), Predicate("ts", "lt", gen.FILTER_HI)))
        if where_tree:  # the same rows as a tree the scan evaluates per row (no fused range)
            where = Not(Or((Predicate("ts", "lt", gen.FILTER_LO), Predicate("ts", "gte", gen.FILTER_HI))))
        return Request(["region"], [AggSpec("count", None, "c"), AggSpec("sum", "amount", "s"),
                                    AggSpec("avg", "price", "a")], [SortSpec("s", True)], 100,
                       where_filter=where, having_filter=Predicate("a", "gt", 500) if having else None)
    if workload == "plain":
        return Request(["k"], [AggSpec("count", None, "c"), AggSpec("sum", "v", "s"), AggSpec("min", "v", "mn"),
                               AggSpec("max", "v", "mx")],
                       having_filter=And((Predicate("c", "gt", 3900), Predicate("mx", "gt", 0))) if having else None)
    if workload == "wide":
        return Request(["a", "b"], [AggSpec("count", None, "c"), AggSpec("sum", "v", "s")], [SortSpec("c", True)], 10)
    return Request(["user_id", "bucket"], [AggSpec("count", None, "c"), AggSpec("sum", "v", "s"),
                                           AggSpec("max", "w", "m")], [SortSpec("c", True)], 10,
                   est_groups=150_000_000,
                   having_filter=And((Predicate("c", "gte", 2), Predicate("s", "gt", 0))) if having else None)


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def stage_from_segments(sess, table, cols, threads: int = 8) -> dict:
    """Writes the table as segment files in host memory (untimed), then
    stages every block the query reads through the segment reader (footer
    parse, zstd decode on the host, encoded body over PCIe, expansion in
    HBM), one segment per worker thread."""
    from concurrent.futures import ThreadPoolExecutor

    from thindb_amd import gen, gpu
    t0 = time.perf_counter()
    imgs = gen.segment_images(table)
    log(f"segment files: {len(imgs)} images, {sum(map(len, imgs.values())) / 1e9:.2f} GB, "
        f"written in {time.perf_counter() - t0:.1f}s")

    def one(img):
        seg = gpu.Segment(img, len(table.schema))
        n = sess.stage_segment(table, seg, cols)
        seg.close()
        return n

    ids = sorted(imgs)
    sync()
    t0 = time.perf_counter()
    blocks = one(imgs[ids[0]])  # the first reserves the table's arenas
    with ThreadPoolExecutor(threads) as ex:
        blocks += sum(ex.map(one, (imgs[i] for i in ids[1:])))
    sync()
    secs = time.perf_counter() - t0
    read = sum(len(imgs[i]) for i in ids)
    log(f"ingested {table.n_rows} rows ({blocks} blocks) from segment files in {secs:.2f}s: "
        f"{table.n_rows / secs / 1e6:.0f} M rows/s, {read / secs / 1e9:.2f} GB/s of file")
    return {"rows": table.n_rows, "blocks": blocks, "seconds": secs, "rows_per_s": table.n_rows / secs,
            "file_bytes": read, "file_gbps": read / secs / 1e9, "threads": threads,
            "codec": "zstd level 3, FOR or raw bodies"}


def host_threads() -> int:
    """Host cores this job may use: its CPU affinity, at most 16 (the box's share per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _time_oracle(table, request, n: int, threads: int, reps: int = 3, radix: bool = False):
    from oracle import oracle
    from thindb_amd.exec import Table
    sample = table if n >= table.n_rows else Table(
        table.schema, {k: v[:n] for k, v in table.columns.items()},
        {k: v[:n] for k, v in table.validity.items()}, rg_rows=table.rg_rows, segment_rows=table.segment_rows,
        uid=table.uid)
    oracle.run(sample, request, threads, radix=radix)  # warm
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.run(sample, request, threads, radix=radix)
        times.append(time.perf_counter() - t0)
    return min(n, table.n_rows), sorted(times)[len(times) // 2], min(times)


def cpu_baseline(table, request, sample_rows: int, threads: int, dop1_rows: int, kind_note: str,
                 radix: bool = False) -> dict:
    """The oracle (thinDB-algorithm CPU restatement) on the first
    `sample_rows` rows of the same table with `threads` workers, and on the
    first `dop1_rows` rows with one (radix: the silo's high-cardinality
    arrangement of it)."""
    n, med, best = _time_oracle(table, request, sample_rows, threads, radix=radix)
    n1, med1, best1 = _time_oracle(table, request, dop1_rows, 1, radix=radix)
    return {"value": n / med, "unit": "rows/s", "cores": threads, "kind": "port",
            "sample": f"first {n} rows of the same table, median of 3 runs after 1 warm-up "
                      f"(best {n / best:.4g} rows/s); {kind_note}",
            "dop1": {"value": n1 / med1, "unit": "rows/s", "cores": 1,
                     "sample": f"first {n1} rows, median of 3 runs after 1 warm-up"}}


ORACLE_NOTE = ("oracle = C restatement of thinDB's grouped aggregate (per-worker tables merged in worker order, "
               "the low-cardinality core's algorithm), not thinDB itself")


_PEAK_CODE = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import bench
print(json.dumps(bench._measured_peak_here()), flush=True)
"""


def measured_peak() -> dict:
    """The device's read and copy bandwidth in this run, measured by torch in
    a child process: torch bundles its own HIP runtime, which cannot share the
    device with the one libthindb_gpu.so already initialised in this process."""
    import subprocess
    r = subprocess.run([sys.executable, "-c", _PEAK_CODE, str(ROOT)], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"bandwidth probe failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def _measured_peak_here() -> dict:
    """torch's int64 sum over 5.6 GB (read) and copy_ of 2.8 GB (read +
    write), HIP events, median of 10 after 3 warm-ups."""
    import torch
    n = 5_600_000_000 // 8
    x = torch.ones(n, dtype=torch.int64, device="cuda")
    y = torch.empty(n // 2, dtype=torch.int64, device="cuda")
    out = {}
    for name, fn, nbytes in (("read_gbps", lambda: x.sum(), n * 8), ("copy_gbps", lambda: y.copy_(x[: n // 2]), n * 8)):
        for _ in range(3):
            fn()
        ms = []
        for _ in range(10):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        out[name] = nbytes / (float(np.median(ms)) / 1e3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    out["method"] = "torch int64 sum over 5.6 GB (read); torch copy_ of 2.8 GB (bytes read + written)"
    return out


def run_batched(args, sess=None, steps: int = 0, warmup: int = 0) -> dict:
    """The batched workload: one batch of all queries vs the same plans run
    one after another; both deliver every result to host memory."""
    import ctypes as C

    from oracle import oracle
    from thindb_amd import gen, gpu
    from thindb_amd.exec import GpuBatchRun, GpuSession

    steps = steps or args.steps
    warmup = warmup or args.warmup
    own = sess is None
    n_tables = args.tables
    t0 = time.perf_counter()
    pairs = [(gen.batched_table(BATCH_SEED, t, args.rows), gen.batched_request(BATCH_SEED, t)) for t in range(n_tables)]
    rows = sum(t.n_rows for t, _ in pairs)
    log(f"generated {n_tables} tables, {rows} rows in {time.perf_counter() - t0:.1f}s")
    if own:
        sess = GpuSession(0)
    t0 = time.perf_counter()
    run = GpuBatchRun(sess, pairs)
    stage_s = time.perf_counter() - t0
    log(f"staged and planned in {stage_s:.1f}s")
    plans = [it[2] for it in run.items]
    lib = gpu.load()

    def sequential() -> None:
        out = C.c_void_p()
        for p in plans:
            gpu.check(lib.tdbg_execute(p.handle, None, 0, C.byref(out)))
            lib.tdbg_result_free(out)

    for _ in range(warmup):
        sequential()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        sequential()
    sync()
    seq_s = (time.perf_counter() - t0) / steps
    log(f"sequential: {seq_s * 1e3:.2f} ms per {n_tables} queries ({n_tables / seq_s:.1f} queries/s)")

    t0 = time.perf_counter()
    run.autotune(2)
    tune_s = time.perf_counter() - t0
    for _ in range(warmup):
        run.batch.execute()
    routes = [run.batch.route(i) for i in range(n_tables)]
    log(f"autotuned in {tune_s:.1f}s; strategies " +
        str({gpu.Batch.STRATEGIES[k]: sum(r.strategy == k for r in routes) for k in gpu.Batch.STRATEGIES}))
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.batch.execute()
    sync()
    batch_s = (time.perf_counter() - t0) / steps
    st = run.batch.stats()
    log(f"batched: {batch_s * 1e3:.2f} ms per {n_tables} queries ({n_tables / batch_s:.1f} queries/s), "
        f"device {st.total_ms:.2f} ms, {st.n_items} items, graph {st.graph}, results {st.result_bytes / 1e6:.1f} MB")

    # the persistent kernel alone: the same plans as an eagerly launched batch
    os.environ["THINDB_GPU_GRAPH"] = "0"  # read when the batch is planned, at its first execution
    eager = gpu.Batch(sess.ctx, plans)
    kern = []
    for _ in range(3):
        eager.execute()
        kern.append(eager.stats().scan_ms)
    del os.environ["THINDB_GPU_GRAPH"]
    eager.close()
    kernel_ms = float(np.median(kern))
    alg_bytes = int(st.bytes)
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    log(f"persistent kernel {kernel_ms:.3f} ms for {alg_bytes / 1e9:.2f} GB ({achieved:.0f} GB/s)")
    # the same kernel over each strategy's plans alone
    breakdown = {}
    for k, name in gpu.Batch.STRATEGIES.items():
        sub = [p for p, r in zip(plans, routes) if r.strategy == k]
        if not sub:
            continue
        os.environ["THINDB_GPU_GRAPH"] = "0"
        b = gpu.Batch(sess.ctx, sub)
        ms = []
        for _ in range(3):
            b.execute()
            ms.append(b.stats().scan_ms)
        del os.environ["THINDB_GPU_GRAPH"]
        bs = b.stats()
        b.close()
        breakdown[name] = {"plans": len(sub), "rows": int(bs.rows), "kernel_ms": float(np.median(ms)),
                           "gbps": bs.bytes / (float(np.median(ms)) / 1e3) / 1e9}
        log(f"  {name}: {len(sub)} plans, {bs.rows / 1e6:.1f}M rows, kernel {np.median(ms):.3f} ms "
            f"({breakdown[name]['gbps']:.0f} GB/s)")

    cpu = None
    if not args.no_cpu:
        threads = host_threads()
        oracle.run(pairs[0][0], pairs[0][1], threads)
        t0 = time.perf_counter()
        for t, r in pairs:
            oracle.run(t, r, threads)
        cpu_s = time.perf_counter() - t0
        # one core: every 16th query (the same uneven mix, a sixteenth of it)
        sub = pairs[::16]
        t0 = time.perf_counter()
        for t, r in sub:
            oracle.run(t, r, 1)
        cpu1_s = time.perf_counter() - t0
        cpu = {"value": n_tables / cpu_s, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"all {n_tables} queries once, after a warm-up query; {ORACLE_NOTE}",
               "dop1": {"value": len(sub) / cpu1_s, "unit": "queries/s", "cores": 1,
                        "sample": f"every 16th query ({len(sub)} of {n_tables}), once"}}
        log(f"cpu baseline {cpu['value']:.2f} queries/s on {threads} cores, {cpu['dop1']['value']:.2f} on one")
    pmc = ROOT / "profiles" / "batch_kernel_512_pmc.json"
    traffic = json.loads(pmc.read_text())["traffic_bytes"] if pmc.exists() and n_tables == 512 and not args.rows \
        else None

    out = {
        "metric": "rows/s & HBM GB/s for filtered GROUP BY; batched QPS + p50/p99 latency",
        "value": n_tables / batch_s,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": batch_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": WORKLOADS["batched"][2], "tables": n_tables, "rows": rows, "seed": BATCH_SEED},
        "sequential": {"value": n_tables / seq_s, "unit": "queries/s", "ms_per_step": seq_s * 1e3},
        "speedup_vs_sequential": seq_s / batch_s,
        "rows_per_s": rows / batch_s,
        "device_ms_per_step": st.total_ms,
        "result_bytes": int(st.result_bytes),
        "autotune_s": tune_s,
        "kernel_by_strategy": breakdown,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "traffic_source": str(pmc.relative_to(ROOT)) if traffic else None, "kernel": "batch_kernel",
                     "kernel_ms": kernel_ms, "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    run.close()
    if own:
        sess.close()
    return out


SERVE_RATES = (100, 250, 500, 1000, 2500, 5000)


def request_bytes(table, req) -> int:
    """Native-width bytes of every column a request reads, over the whole table."""
    from thindb_amd.gpu import NUMPY_DTYPE
    names = {c for c in req.group_cols}
    names |= {a.col for a in req.aggs if a.col}
    stack = [req.where_filter] if req.where_filter is not None else []
    while stack:
        e = stack.pop()
        if hasattr(e, "children"):
            stack.extend(e.children)
        elif hasattr(e, "child"):
            stack.append(e.child)
        elif hasattr(e, "col"):
            names.add(e.col)
    return sum(np.dtype(NUMPY_DTYPE[table.column(n).type]).itemsize for n in names) * table.n_rows


def run_serving(args, sess=None, big=None) -> dict:
    """The serving loop under Poisson arrivals: 70% small (batched-workload
    tables of 1e4-1e6 rows), 25% medium (1e7 rows, 1e5 groups), 5% large (the
    200M-row filtered query).  The served form (tdbg_server: continuous
    batching, 3 lanes, admission by estimated bytes) against the plain
    operator behind a FIFO queue (one lane, no batching), both through the
    same C-ABI; latency = completion - arrival (steady clock)."""
    import threading

    from thindb_amd import gen, gpu
    from thindb_amd.exec import And, AggSpec, GpuSession, Predicate, Request, SortSpec, build_plan

    rng = np.random.default_rng(41)
    own = sess is None
    if own:
        sess = GpuSession(0)
    t0 = time.perf_counter()
    classes = {"small": [], "medium": [], "large": []}
    for t in range(48):
        n = int(round(np.exp(rng.uniform(np.log(1e4), np.log(1e6)))))
        classes["small"].append((gen.batched_table(41, t, n), gen.batched_request(41, t)))
    for t in range(6):
        classes["medium"].append((gen.keyed_table(100 + t, 10_000_000, 100_000), gen.batched_request(41, t)))
    if big is None:
        big = gen.table("filtered", args.rows or 200_000_000)
    classes["large"].append((big, make_request("filtered")))
    plans = {k: [] for k in classes}
    staged = 0
    plan_bytes = {}
    plan_class = {}
    for k, pairs in classes.items():
        for table, req in pairs:
            plan, cols = build_plan(sess, table, req)
            staged += sess.stage_table(table, cols)
            plans[k].append(plan)
            plan_bytes[id(plan)] = request_bytes(table, req)
            plan_class[id(plan)] = k
    log(f"resident: {staged / 1e9:.2f} GB in {sum(len(v) for v in plans.values())} tables, "
        f"{time.perf_counter() - t0:.1f}s")

    def draw(n, seed):
        r = np.random.default_rng(seed)
        u = r.random(n)
        out = []
        for x in u:
            k = "small" if x < 0.70 else ("medium" if x < 0.95 else "large")
            out.append(plans[k][r.integers(len(plans[k]))])
        return out

    # single-query latency of each class, alone on an idle GPU
    # (submission to completion through the plain operator, one request at a time)
    single = {}
    solo = gpu.Server(sess.ctx, lanes=1, max_batch=1)
    for k, ps in plans.items():
        lat = []
        for p in ps[:8]:
            for rep in range(4):
                t = solo.submit(p)
                t.wait()
                r = t.report()
                t.close()
                if rep:
                    lat.append((r.done_ns - r.submit_ns) / 1e9)
        single[k] = float(np.median(lat)) * 1e3
    solo.stop()
    mix_single = 0.70 * single["small"] + 0.25 * single["medium"] + 0.05 * single["large"]
    log("single-query latency ms: " + ", ".join(f"{k} {v:.3f}" for k, v in single.items()) +
        f"; mix-weighted {mix_single:.3f}")

    def drive(server, reqs, rate):
        """Submits reqs at Poisson arrival times (rate/s; 0 = all at once)."""
        r = np.random.default_rng(int(rate) + 7)
        gaps = r.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
        times = np.cumsum(gaps)
        tickets = []
        start = time.perf_counter()
        for p, at in zip(reqs, times):
            while time.perf_counter() - start < at:
                pass
            tickets.append(server.submit(p))
        cls = [plan_class[id(p)] for p in reqs]
        served_bytes = sum(plan_bytes[id(p)] for p in reqs)
        lat, first, last = [], None, None
        units = {}
        for t in tickets:
            t.wait()
            rep = t.report()
            units[(rep.lane, rep.start_ns)] = (rep.batch_size, (rep.done_ns - rep.start_ns) / 1e6)
            lat.append((rep.done_ns - rep.submit_ns) / 1e6)
            first = rep.submit_ns if first is None else min(first, rep.submit_ns)
            last = rep.done_ns if last is None else max(last, rep.done_ns)
            t.close()
        lat = np.array(lat)
        # per request class: p50/p99 and the ratio of p99 to the class's single-query latency
        per_class = {}
        for k in ("small", "medium", "large"):
            lk = np.array([x for x, c in zip(lat, cls) if c == k])
            if len(lk):
                per_class[k] = {"n": int(len(lk)), "p50_ms": float(np.percentile(lk, 50)),
                                "p99_ms": float(np.percentile(lk, 99)),
                                "p99_over_single": float(np.percentile(lk, 99)) / single[k]}
        # launches by size: how many, mean duration (ms)
        by = {}
        for size, ms in units.values():
            k = "1" if size == 1 else ("2-8" if size <= 8 else ("9-32" if size <= 32 else "33+"))
            n, tot = by.get(k, (0, 0.0))
            by[k] = (n + 1, tot + ms)
        return {"offered_qps": rate, "throughput_qps": len(reqs) / ((last - first) / 1e9),
                "served_gbps": served_bytes / ((last - first) / 1e9) / 1e9,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "per_class": per_class,
                "launches": {k: {"count": n, "mean_ms": tot / n} for k, (n, tot) in sorted(by.items())}}

    results = {"served": [], "plain": []}
    for name, cfg in (("served", dict(lanes=3, max_batch=64)), ("plain", dict(lanes=1, max_batch=1))):
        srv = gpu.Server(sess.ctx, **cfg)
        drive(srv, draw(200, 1), 0)  # warm-up: plans, buffers, kernels
        for rate in SERVE_RATES:
            n = int(min(max(rate * args.serve_seconds, 200), 6000))
            res = drive(srv, draw(n, rate), rate)
            res["requests"] = n
            results[name].append(res)
            log(f"{name} {rate}/s: {res['throughput_qps']:.0f} q/s, p50 {res['p50_ms']:.2f} ms, "
                f"p99 {res['p99_ms']:.2f} ms; p99 / single by class: "
                + ", ".join(f"{k} {v['p99_over_single']:.1f}" for k, v in res["per_class"].items()))
        sat = drive(srv, draw(3000, 99), 0)
        sat["requests"] = 3000
        results[name + "_saturated"] = sat
        c = srv.counters()
        log(f"{name} saturated: {sat['throughput_qps']:.0f} q/s, p50 {sat['p50_ms']:.2f} ms, p99 {sat['p99_ms']:.2f} ms; "
            f"units {c.units}, batched {c.batched}, admission waits {c.admission_waits}; launches {sat['launches']}")
        srv.stop()
    served, plain = results["served_saturated"], results["plain_saturated"]
    # the CPU path over the same mix: the oracle's mean time per request of
    # each class (a sample of the class's requests; the large query on a
    # 40M-row prefix, scaled by rows), weighted 70/25/5, closed loop
    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        threads = host_threads()
        per = {}
        for k, pairs in classes.items():
            ts = []
            for table, req in pairs[:8]:
                n = min(table.n_rows, 40_000_000)
                got_n, med, _ = _time_oracle(table, req, n, threads, reps=1)
                ts.append(med * table.n_rows / got_n)
            per[k] = float(np.mean(ts))
        mix_s = 0.70 * per["small"] + 0.25 * per["medium"] + 0.05 * per["large"]
        cpu = {"value": 1.0 / mix_s, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": "mean oracle time per request of each class (up to 8 requests per class; the large "
                         "query timed on its first 40M rows and scaled by rows), weighted 70/25/5; " + ORACLE_NOTE,
               "per_class_ms": {k: v * 1e3 for k, v in per.items()}}
        log(f"cpu baseline {cpu['value']:.1f} queries/s on {threads} cores")
    out = {
        "metric": "rows/s & HBM GB/s for filtered GROUP BY; batched QPS + p50/p99 latency",
        "value": served["throughput_qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": 1e3 / served["throughput_qps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": "serving: Poisson arrivals, 70% small / 25% medium / 5% large (200M-row filtered)",
                   "resident_gb": staged / 1e9, "rates": list(SERVE_RATES)},
        "single_query_ms": single,
        "sweep": results,
        "speedup_vs_plain_saturated": served["throughput_qps"] / plain["throughput_qps"],
        # the served work at saturation: the algorithmic column bytes of every
        # completed request per second of wall clock (GPU busy or not)
        "roofline": {"bound": "hbm", "achieved": served["served_gbps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": served["served_gbps"] / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "every launch of the served mix (wall clock at saturation)"},
        "cpu_baseline": cpu,
    }
    for ps in plans.values():
        for p in ps:
            p.close()
    if own:
        sess.close()
    return out


def run_single(args, workload: str, sess, steps: int, warmup: int, table=None):
    """One query of `workload` (filtered / plain / highcard) over its
    HBM-resident table; returns (record, table)."""
    from thindb_amd import gen
    from thindb_amd.exec import try_build_gpu_group

    rows, bytes_per_row, desc = WORKLOADS[workload]
    if args.rows and workload == args.workload:
        rows = args.rows
    if table is None:
        log(f"generating {rows} rows ({workload})")
        t0 = time.perf_counter()
        table = gen.table(workload, rows)
        log(f"generated in {time.perf_counter() - t0:.1f}s")
    opts = args.workload == workload  # --having / --distinct / --where-tree apply to the chosen workload
    having, distinct, where_tree = opts and args.having, opts and args.distinct, opts and args.where_tree
    req = make_request(workload, having, distinct, where_tree, args.distinct_col)
    if where_tree:
        desc = f"{desc}, WHERE NOT (ts < lo OR ts >= hi)"
    if having:
        desc = f"{desc} {HAVING[workload]}"
    if distinct:
        desc, bytes_per_row = DISTINCT_DESC.format(col=args.distinct_col), DISTINCT_BYTES_PER_ROW
    q = try_build_gpu_group(sess, table, req)
    if q is None:
        raise RuntimeError(f"GPU operator declined the {workload} workload")
    cols = q._columns
    ingest = None
    if opts and args.from_segments:
        ingest = stage_from_segments(sess, table, cols)
    t0 = time.perf_counter()
    staged = sess.stage_table(table, cols)
    stage_s = time.perf_counter() - t0
    if staged:
        log(f"staged {staged / 1e9:.2f} GB in {stage_s:.2f}s ({staged / stage_s / 1e9:.1f} GB/s incl. PCIe)")

    rgs = q.plan.rg_array([(rg.segment_id, rg.idx) for rg in table.row_groups])
    for _ in range(warmup):
        q.plan.execute(rgs)
    # a step is what the operator's caller does through the C-ABI: execute,
    # read the typed result columns in pinned host memory, free them (the
    # numpy conversion of the test harness is not part of the operator)
    import ctypes as C
    from thindb_amd import gpu
    lib = gpu.load()
    arr, n_rgs = rgs
    res = C.c_void_p()
    scan_ms, total_ms = [], []
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        gpu.check(lib.tdbg_execute(q.plan.handle, arr, n_rgs, C.byref(res)))
        lib.tdbg_result_free(res)
        st = q.plan.stats()
        scan_ms.append(st.scan_ms)
        total_ms.append(st.total_ms)
    sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / steps
    st = q.plan.stats()
    n_rows_scanned = int(st.rows_scanned)
    alg_bytes = n_rows_scanned * bytes_per_row
    if workload == "highcard" and not distinct:
        # plus the compulsory group state: key 9 + count 8 + sum 8 + max 8 bytes per group
        alg_bytes += int(st.n_groups) * HIGHCARD_GROUP_BYTES
    if workload == "wide":
        alg_bytes += int(st.n_groups) * WIDE_GROUP_BYTES
    if distinct:  # the distinct pass runs after the scan: price the whole device time
        scan_ms = total_ms
    avg_scan_s = float(np.mean(scan_ms)) / 1e3
    achieved = alg_bytes / avg_scan_s / 1e9
    log(f"{workload}: {ms_per_step:.3f} ms/step wall, scan kernel {np.mean(scan_ms):.3f} ms avg "
        f"(min {np.min(scan_ms):.3f}), device total {np.mean(total_ms):.3f} ms, groups {st.n_groups}, "
        f"lds_mode {st.lds_mode} global_mode {st.global_mode}")

    cpu = None
    if not args.no_cpu:
        threads = host_threads()
        big = workload in ("filtered", "highcard", "wide")
        hc = workload in ("highcard", "wide") and not distinct
        sample = args.cpu_sample_rows or min(rows, {"filtered": 40_000_000, "wide": 20_000_000,
                                                    "highcard": 20_000_000}.get(workload, 10_000_000))
        cpu = cpu_baseline(table, req, sample, threads,
                           min(sample, 4_000_000 if big else rows),
                           (ORACLE_NOTE + "; for this high-cardinality query in the silo's arrangement "
                            "(oracle_group_by_radix: rows routed by key hash into 256 buckets, each split into "
                            "parts of a few thousand rows folded into tables of their own, silo_group_core.zig)")
                           if hc else ORACLE_NOTE, radix=hc)
        log(f"cpu baseline {cpu['value']:.4g} rows/s on {threads} cores, {cpu['dop1']['value']:.4g} on one")

    # HBM bytes per launch of the dominant kernel, from the committed
    # FETCH_SIZE/WRITE_SIZE counter passes (scaled by a 1 GiB copy kernel).
    traffic = None
    radix = int(st.global_mode) in (3, 4)
    size = f"{WORKLOADS[workload][0] // 1_000_000}M"
    pmc = ROOT / "profiles" / (f"radix_{workload}_{size}_pmc.json" if radix else f"scan_jit_{workload}_{size}_pmc.json")
    if pmc.exists() and rows == WORKLOADS[workload][0] and (st.specialized or radix) and not having \
            and not distinct and not where_tree:
        traffic = json.loads(pmc.read_text())["traffic_bytes"]

    out = {
        "metric": "rows/s & HBM GB/s for filtered GROUP BY; batched QPS + p50/p99 latency",
        "value": n_rows_scanned / (ms_per_step / 1e3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": desc, "rows": n_rows_scanned, "row_group_rows": table.rg_rows,
                   "segment_rows": table.segment_rows, "groups": int(st.n_groups)},
        "hbm_gbps": alg_bytes / (ms_per_step / 1e3) / 1e9,
        "device_ms_per_step": float(np.mean(total_ms)),
        "stage_gbps_pcie": staged / stage_s / 1e9 if staged else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "traffic_source": str(pmc.relative_to(ROOT)) if traffic else None,
                     "kernel": ("device total: scan + distinct_insert_kernel + gather" if distinct
                                else "hash-partition pipeline rx_hist+rx_scatter1+rx_hist2+rx_scatter2+rx_agg_hashed"
                                if int(st.global_mode) == 4
                                else "radix pipeline rx_hist+rx_scatter1+rx_hist2+rx_scatter2+rx_agg" if radix
                                else "tdbg_scan_jit" if st.specialized else "scan_agg_kernel"),
                     "kernel_ms": float(np.mean(scan_ms)),
                     "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        "ingest": ingest,
    }
    q.deinit()
    return out, table


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all", choices=["all"] + sorted(WORKLOADS))
    ap.add_argument("--rows", type=int, default=0,
                    help="override the row count (batched: cap per table; testing only)")
    ap.add_argument("--tables", type=int, default=512, help="batched: number of tables/queries")
    ap.add_argument("--serve-seconds", type=float, default=1.5, help="serving: seconds of arrivals per rate")
    ap.add_argument("--having", action="store_true",
                    help="filtered/plain/highcard: add the workload's HAVING (bench.HAVING)")
    ap.add_argument("--distinct", action="store_true",
                    help="highcard: GROUP BY bucket, COUNT(DISTINCT user_id) instead of the headline query")
    ap.add_argument("--distinct-col", default="user_id", choices=["user_id", "w"],
                    help="--distinct over user_id (BIGINT) or w (DOUBLE: the wide, row-entry form)")
    ap.add_argument("--where-tree", action="store_true",
                    help="filtered: the WHERE as NOT (ts < lo OR ts >= hi), evaluated as a tree")
    ap.add_argument("--from-segments", action="store_true",
                    help="stage the table from thinDB segment files (zstd, FOR/raw blocks) and report ingestion")
    ap.add_argument("--cpu-sample-rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a device-to-device copy of MIB MiB (a known-size kernel that scales "
                         "the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this engine has no multi-GPU path (asked for {args.gpus}); "
              "see the sharding design in DESIGN.md", file=sys.stderr)
        return 2
    if args.distinct and args.workload != "highcard":
        log("--distinct runs over the highcard workload")
        return 2

    from thindb_amd.exec import GpuSession

    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n // 8, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 copies of {n} bytes")
        del src, dst

    if args.workload == "batched":
        out = run_batched(args)
    elif args.workload == "serving":
        out = run_serving(args)
    elif args.workload != "all":
        sess = GpuSession(0)
        out, _ = run_single(args, args.workload, sess, args.steps, args.warmup)
        sess.close()
    else:
        # every configuration in one record: the headline at the top level
        t_start = time.perf_counter()
        sess = GpuSession(0)
        out, filtered = run_single(args, "filtered", sess, args.steps, args.warmup)
        peak = measured_peak()  # after the headline, so nothing of torch's runs before it
        log(f"measured peak: read {peak['read_gbps']:.0f} GB/s, copy {peak['copy_gbps']:.0f} GB/s")
        out["roofline"]["frac_of_measured_read"] = out["roofline"]["achieved"] / peak["read_gbps"]
        configs = {}
        configs["plain"], _ = run_single(args, "plain", sess, max(args.steps, 20), max(args.warmup, 3))
        hc, hc_table = run_single(args, "highcard", sess, min(args.steps, 10), min(args.warmup, 2))
        configs["highcard"] = hc
        sess.evict_table(hc_table)
        del hc_table
        wd, wd_table = run_single(args, "wide", sess, min(args.steps, 10), min(args.warmup, 2))
        configs["wide"] = wd
        sess.evict_table(wd_table)
        del wd_table
        configs["batched"] = run_batched(args, sess, steps=min(args.steps, 5), warmup=min(args.warmup, 2))
        # the serving loop in a context of its own, as a serving process has
        # (a context keeps the strategies the batched run autotuned on its
        # own 512 plans, which the serving loop's batches would then take)
        configs["serving"] = run_serving(args, None, big=filtered)
        for c in configs.values():
            if c.get("roofline"):
                c["roofline"]["frac_of_measured_read"] = c["roofline"]["achieved"] / peak["read_gbps"]
        out["configs"] = configs
        out["measured_peak"] = peak
        out["bench_seconds"] = time.perf_counter() - t_start
        sess.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
