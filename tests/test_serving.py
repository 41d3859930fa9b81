"""The device request loop (tdbg_server_*), through the C-ABI: every request's
result equals the oracle's (bit-exact integers, DOUBLE within 1e-12
relative) and is the same alone and in a concurrent mix that the server
batches, overlaps on its lanes and admits by estimated device bytes."""
import numpy as np
import pytest

from helpers import assert_same_rows, check_with_oracle, float_cols_of, order_cols_of, rows_of
from thindb_amd import gen, gpu
from thindb_amd.exec import And, AggSpec, Batch, Predicate, Request, SortSpec, build_plan, output_schema

pytestmark = pytest.mark.gpu

REL = 1e-12


def mix(session):
    """Small batched-workload tables, two medium tables and one large filtered
    query with ORDER BY ... LIMIT (runs alone), with their plans."""
    items = []
    for t in range(24):
        table = gen.batched_table(21, t, 60_000)
        items.append((table, gen.batched_request(21, t)))
    for t in (2, 5):
        table = gen.batched_table(22, t, 2_000_000)
        items.append((table, gen.batched_request(22, t)))
    big = gen.table("filtered", 3_000_000, seed=9)
    where = And((Predicate("ts", "gte", gen.FILTER_LO), Predicate("ts", "lt", gen.FILTER_HI)))
    items.append((big, Request(["region"], [AggSpec("count", None, "c"), AggSpec("sum", "amount", "s"),
                                            AggSpec("avg", "price", "a")], [SortSpec("s", True)], 100,
                               where_filter=where)))
    plans = []
    for table, req in items:
        plan, cols = build_plan(session, table, req)
        session.stage_table(table, cols)
        plans.append(plan)
    return items, plans


def alone(items, plans):
    out = []
    for (table, req), p in zip(items, plans):
        cols = p.execute()
        out.append(rows_of(Batch(output_schema(table, req), cols, len(cols[0].values))))
    return out


def compare(items, got, want):
    """Served results against the same requests run alone: identical rows."""
    for (table, req), g, w in zip(items, got, want):
        names = [c.name for c in output_schema(table, req)]
        assert_same_rows(g, w, float_cols_of(table, req), REL, len(req.group_cols), order_cols_of(names, req),
                         full=w if req.order_specs else None)


def oracle_rows(oracle_lib, items):
    return [rows_of(oracle_lib.run(table, req, 8)) for table, req in items]


def compare_oracle(oracle_lib, items, got, want_oracle):
    """Served results against the oracle (the CPU path) for each request."""
    for (table, req), g, w in zip(items, got, want_oracle):
        check_with_oracle(g, oracle_lib, table, req, 8, REL, want=w)


def test_same_alone_and_in_the_mix(session, oracle_lib):
    items, plans = mix(session)
    want = alone(items, plans)
    want_oracle = oracle_rows(oracle_lib, items)
    compare_oracle(oracle_lib, items, want, want_oracle)
    print(f"[gpu] serving mix of {len(items)} requests, 3 rounds", flush=True)
    srv = gpu.Server(session.ctx, lanes=3, max_batch=16)
    try:
        for rnd in range(3):
            order = np.random.default_rng(rnd).permutation(len(items))
            tickets = [(i, srv.submit(plans[i])) for i in order]  # all at once: queued behind running units
            got = [None] * len(items)
            sizes = []
            for i, t in tickets:
                cols = t.wait()
                table, req = items[i]
                got[i] = rows_of(Batch(output_schema(table, req), cols, len(cols[0].values)))
                rep = t.report()
                assert rep.status == 0 and rep.done_ns >= rep.start_ns >= rep.submit_ns
                sizes.append(rep.batch_size)
                t.close()
            compare(items, got, want)
            compare_oracle(oracle_lib, items, got, want_oracle)
            print(f"  round {rnd}: batch sizes {sorted(set(sizes))}", flush=True)
        c = srv.counters()
        assert c.completed == c.submitted == 3 * len(items)
        assert c.batched > 0, "queued small requests share launches"
    finally:
        srv.stop()
        for p in plans:
            p.close()


def test_admission_by_device_bytes(session, oracle_lib):
    items, plans = mix(session)
    want = alone(items, plans)
    want_oracle = oracle_rows(oracle_lib, items)
    # room for about one medium request at a time: later arrivals wait for memory
    est = []
    srv = gpu.Server(session.ctx, lanes=3, max_batch=8, admit_bytes=64 << 20)
    try:
        tickets = [srv.submit(p) for p in plans]
        got = []
        for (table, req), t in zip(items, tickets):
            cols = t.wait()
            est.append(t.report().est_bytes)
            got.append(rows_of(Batch(output_schema(table, req), cols, len(cols[0].values))))
            t.close()
        compare(items, got, want)
        compare_oracle(oracle_lib, items, got, want_oracle)
        c = srv.counters()
        print(f"[gpu] admission: peak {c.peak_admitted >> 20} MiB of 64, waits {c.admission_waits}, "
              f"largest estimate {max(est) >> 20} MiB", flush=True)
        assert c.admission_waits > 0
        assert c.peak_admitted <= max(64 << 20, max(est))
    finally:
        srv.stop()
        for p in plans:
            p.close()


def test_requests_over_named_row_groups(session, oracle_lib):
    """tdbg_server_submit_rgs: each request scans the row groups it names (a
    query's snapshot), batched or alone, and equals the plan executed over
    the same row groups and the oracle over those rows; an empty list scans
    nothing."""
    from thindb_amd.exec import Table
    items, plans = [], []
    for t in range(12):  # small tables of 8K-row row groups: batched, over a subset of them
        base = gen.batched_table(23, t, 60_000)
        table = Table(base.schema, base.columns, base.validity, rg_rows=8192)
        items.append((table, gen.batched_request(23, t)))
    big = gen.table("filtered", 3_000_000, seed=10)
    where = And((Predicate("ts", "gte", gen.FILTER_LO), Predicate("ts", "lt", gen.FILTER_HI)))
    items.append((big, Request(["region"], [AggSpec("count", None, "c"), AggSpec("sum", "amount", "s")],
                               [SortSpec("s", True)], 50, where_filter=where)))
    for table, req in items:
        plan, cols = build_plan(session, table, req)
        session.stage_table(table, cols)
        plans.append(plan)
    picked = []
    for (table, req), p in zip(items, plans):
        rgs = table.row_groups
        keep = [rg for i, rg in enumerate(rgs) if i % 3 != 1] or rgs[:1]
        picked.append(keep)
    want = []
    for (table, req), p, keep in zip(items, plans, picked):
        cols = p.execute([(rg.segment_id, rg.idx) for rg in keep])
        want.append(rows_of(Batch(output_schema(table, req), cols, len(cols[0].values) if cols else 0)))
    srv = gpu.Server(session.ctx, lanes=3, max_batch=16)
    try:
        tickets = [srv.submit(p, [(rg.segment_id, rg.idx) for rg in keep]) for p, keep in zip(plans, picked)]
        got = []
        sizes = []
        for (table, req), t in zip(items, tickets):
            cols = t.wait()
            got.append(rows_of(Batch(output_schema(table, req), cols, len(cols[0].values) if cols else 0)))
            sizes.append(t.report().batch_size)
            t.close()
        compare(items, got, want)
        assert max(sizes) > 1, "requests over named row groups share launches"
        # the oracle over exactly the picked rows
        for (table, req), g, keep in zip(items, got, picked):
            idx = np.concatenate([np.arange(rg.start, rg.start + rg.n_rows) for rg in keep])
            sub = Table(table.schema, {k: v[idx] for k, v in table.columns.items()},
                        {k: v[idx] for k, v in table.validity.items()}, rg_rows=table.rg_rows)
            check_with_oracle(g, oracle_lib, sub, req, 8, REL)
        t = srv.submit(plans[0], [])
        cols = t.wait()
        t.close()
        assert all(len(c.values) == 0 for c in cols)
    finally:
        srv.stop()
        for p in plans:
            p.close()
