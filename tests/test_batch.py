"""Batched execution: many plans in one persistent launch (tdbg_batch_*),
called through the C-ABI.  Every plan's result must equal the oracle's (and
the plan's own single-query result) whatever strategy the batch routes it
through, on every replay of the captured graph."""
import numpy as np
import pytest

from helpers import ALL_AGGS, assert_same_rows, random_table, rows_of
from thindb_amd import gen
from thindb_amd.exec import (And, AggSpec, Declined, GpuBatchRun, Predicate, Request, SortSpec,
                             try_build_gpu_group)

pytestmark = pytest.mark.gpu

REL = 1e-12


def float_cols(table, req):
    from test_gpu_parity import float_cols_of
    return float_cols_of(table, req)


def batched_pairs(seed, n, max_rows):
    return [(gen.batched_table(seed, t, max_rows), gen.batched_request(seed, t)) for t in range(n)]


def typed_pairs():
    """Mixed typed tables (NULLs, every aggregate kind, filters)."""
    pairs = []
    where = And((Predicate("i", "gt", -2 ** 30), Predicate("s", "neq", 0)))
    for seed in range(6):
        rng = np.random.default_rng(100 + seed)
        t = random_table(rng, 20_000 + 7_919 * seed, rg_rows=3_001 + seed)
        pairs.append((t, Request(["k1"], ALL_AGGS)))
        pairs.append((t, Request(["k1", "k2"], ALL_AGGS, where_filter=where)))
        pairs.append((t, Request(["k2"], ALL_AGGS[:7], where_filter=Predicate("d", "lt", 1000.0))))
    return pairs


def check(oracle_lib, pairs, results):
    for (t, r), b in zip(pairs, results):
        want = rows_of(oracle_lib.run(t, r, 4))
        assert_same_rows(rows_of(b), want, float_cols(t, r), REL, len(r.group_cols), None)


def run_batch(session, pairs, tune=False):
    print(f"[gpu] batch of {len(pairs)} plans, {sum(t.n_rows for t, _ in pairs)} rows tune={tune}", flush=True)
    run = GpuBatchRun(session, pairs)
    try:
        if tune:
            run.autotune(1)
        res = run.execute()
        routes = [run.batch.route(i) for i in range(len(pairs))]
        st = run.batch.stats()
    finally:
        run.close()
    return res, routes, st


def test_batched_workload_vs_oracle(session, oracle_lib):
    pairs = batched_pairs(7, 48, 150_000)
    res, routes, st = run_batch(session, pairs)
    check(oracle_lib, pairs, res)
    assert st.n_plans == 48 and st.graph == 1
    print("  strategies:", sorted({r.strategy for r in routes}), f"items {st.n_items} {st.total_ms:.3f} ms")


@pytest.mark.parametrize("strategy", ["direct", "global", "hash", "radix"])
def test_each_strategy_vs_oracle(session, oracle_lib, monkeypatch, strategy):
    monkeypatch.setenv("THINDB_GPU_BATCH_STRATEGY", strategy)
    pairs = typed_pairs() + batched_pairs(3, 10, 100_000)
    res, routes, _ = run_batch(session, pairs)
    check(oracle_lib, pairs, res)
    want = {"direct": 1, "global": 2, "hash": 3, "radix": 4}[strategy]
    taken = sum(r.strategy == want for r in routes)
    print(f"  {strategy}: {taken} of {len(routes)} plans")
    assert taken > 0


def test_same_as_alone(session):
    """Each plan's batched result is its single-query result: integer columns
    bit-exact, DOUBLE SUM/AVG within 1e-12 (their addition order differs)."""
    pairs = typed_pairs()[:6] + batched_pairs(5, 6, 80_000)
    res, _, _ = run_batch(session, pairs)
    for (t, r), b in zip(pairs, res):
        q = try_build_gpu_group(session, t, r)
        alone = rows_of(q.next())
        q.deinit()
        assert_same_rows(rows_of(b), alone, float_cols(t, r), REL, len(r.group_cols), None)


def test_graph_replays(session, oracle_lib, monkeypatch):
    """The captured graph replayed several times, with other GPU work in
    between, gives the eager launch sequence's results every time."""
    pairs = batched_pairs(11, 24, 120_000) + typed_pairs()[:3]
    monkeypatch.setenv("THINDB_GPU_GRAPH", "0")
    eager, _, st = run_batch(session, pairs)
    assert st.graph == 0
    monkeypatch.delenv("THINDB_GPU_GRAPH")
    run = GpuBatchRun(session, pairs)
    try:
        for rep in range(3):
            res = run.execute()
            assert run.batch.stats().graph == 1
            for (t, r), a, b in zip(pairs, res, eager):  # DOUBLE sums add in run-dependent order
                assert_same_rows(rows_of(a), rows_of(b), float_cols(t, r), REL, len(r.group_cols), None)
            q = try_build_gpu_group(session, *typed_pairs()[1])  # other work between replays
            q.next()
            q.deinit()
    finally:
        run.close()
    check(oracle_lib, pairs, eager)


def test_autotune_keeps_results(session, oracle_lib):
    pairs = batched_pairs(13, 24, 200_000)
    res, routes, st = run_batch(session, pairs, tune=True)
    check(oracle_lib, pairs, res)
    print("  tuned strategies:", [r.strategy for r in routes])


def test_declines_ordered_requests(session):
    t = gen.batched_table(7, 0, 10_000)
    r = Request(["key"], [AggSpec("count", None, "c")], [SortSpec("c", True)], 5)
    with pytest.raises(Declined):
        GpuBatchRun(session, [(t, r)])


def test_waves_same_results(session, oracle_lib, monkeypatch):
    """A batch large enough to run as waves (result copies overlapping the
    next wave's kernels) gives every plan the result of the one-launch batch
    and of the oracle, on two executions in a row."""
    pairs = batched_pairs(17, 96, 30_000)
    monkeypatch.setenv("THINDB_GPU_BATCH_WAVES", "1")
    single, _, st1 = run_batch(session, pairs)
    assert st1.waves == 0
    monkeypatch.setenv("THINDB_GPU_BATCH_WAVES", "3")
    run = GpuBatchRun(session, pairs)
    try:
        for rep in range(2):
            res = run.execute()
            st = run.batch.stats()
            assert st.waves == 3 and st.n_plans == 96 and st.graph == 1
            assert [run.batch.route(i).items > 0 for i in range(len(pairs))] == [True] * len(pairs)
            for (t, r), a, b in zip(pairs, res, single):
                assert_same_rows(rows_of(a), rows_of(b), float_cols(t, r), REL, len(r.group_cols), None)
    finally:
        run.close()
    check(oracle_lib, pairs[::8], res[::8])


@pytest.mark.slow
def test_batched_full_size_properties(session):
    """The batched config at full size: the 512 uneven tables of the bench
    (1.29G rows) in one batch; per plan, sum of COUNT(*) = the rows passing
    its filter, sum of SUM(value) = their exact total, the group count = the
    distinct keys among them, MIN/MAX = the extremes."""
    seed = 7
    pairs = [(gen.batched_table(seed, t), gen.batched_request(seed, t)) for t in range(512)]
    res, routes, st = run_batch(session, pairs)
    print(f"[gpu] full-size batch: {st.rows} rows, {st.n_items} items, strategies "
          f"{sorted(set(r.strategy for r in routes))}", flush=True)
    for i, ((t, r), b) in enumerate(zip(pairs, res)):
        if i % 64 == 0:
            print(f"  checking plan {i}", flush=True)
        key, val = t.columns["key"], t.columns["value"]
        if r.where_filter is not None:
            m = (val >= gen.BATCH_FILTER_LO) & (val < gen.BATCH_FILTER_HI)
            key, val = key[m], val[m]
        c, s_, mn, mx = (b.values[i].values for i in range(1, 5))
        assert b.row_count == np.unique(key).size
        assert int(c.sum()) == key.size
        assert int(s_.sum()) == int(val.sum())
        if key.size:
            assert int(mn.min()) == int(val.min()) and int(mx.max()) == int(val.max())
