"""tdbg_execute_noalloc (the allocation-free entry point) against tdbg_execute,
through the C-ABI: the same rows (integers bit-exact, DOUBLE AVG within 1e-12
relative) on two calls in a row, the second over a smaller input, for an
ORDER BY ... LIMIT plan over the dense table, a HAVING plan over the HBM hash
table, a radix-partitioned plan and a COUNT(DISTINCT) plan; limits refused
with clear errors."""
import numpy as np
import pytest

from helpers import ALL_AGGS, assert_same_rows, float_cols_of, order_cols_of, random_table, rows_of
from thindb_amd import gen, gpu
from thindb_amd.exec import And, AggSpec, Batch, Predicate, Request, SortSpec, build_plan, output_schema

pytestmark = pytest.mark.gpu

REL = 1e-12
CAP = 8 << 20


def rows(table, req, cols):
    return rows_of(Batch(output_schema(table, req), cols, len(cols[0].values) if cols else 0))


def compare(table, req, got, want):
    names = [c.name for c in output_schema(table, req)]
    assert_same_rows(got, want, float_cols_of(table, req), REL, len(req.group_cols), order_cols_of(names, req),
                     full=want if (req.order_specs or req.limit) else None)


def run_both(session, table, req, rgs_small):
    plan, cols = build_plan(session, table, req)
    session.stage_table(table, cols)
    try:
        for rgs in (None, rgs_small):  # two calls in a row, the second over fewer rows
            print(f"[noalloc] {table.n_rows} rows keys={list(req.group_cols)} rgs={'all' if rgs is None else len(rgs)}",
                  flush=True)
            want = rows(table, req, plan.execute(rgs))
            got = rows(table, req, plan.execute_noalloc(CAP, rgs))
            assert len(got) == len(want)
            compare(table, req, got, want)
        return plan.stats()
    finally:
        plan.close()


def test_noalloc_matches_execute(session):
    t = gen.table("filtered", 2_000_000, seed=3)
    where = And((Predicate("ts", "gte", gen.FILTER_LO), Predicate("ts", "lt", gen.FILTER_HI)))
    req = Request(["region"], [AggSpec("count", None, "c"), AggSpec("sum", "amount", "s"),
                               AggSpec("avg", "price", "a")], [SortSpec("s", True)], 100, where_filter=where)
    small = [(rg.segment_id, rg.idx) for rg in t.row_groups][:7]
    run_both(session, t, req, small)


def test_noalloc_hash_having_radix_distinct(session, monkeypatch):
    rng = np.random.default_rng(8)
    t = random_table(rng, 60_000, rg_rows=4096, nulls=True)
    small = [(rg.segment_id, rg.idx) for rg in t.row_groups][:5]
    # a sparse DATETIME key (60K values over 1e12): the HBM hash table, with HAVING and OFFSET
    run_both(session, t, Request(["t"], ALL_AGGS[:8], [SortSpec("c", True), SortSpec("sb")], 50, 3,
                                 having_filter=Predicate("c", "gte", 1)), small)
    # COUNT(DISTINCT) next to the other aggregates
    run_both(session, t, Request(["k1", "k2"], ALL_AGGS[:5] + [AggSpec("count_distinct", "s", "ds")]), small)
    # radix partitioning forced: compacted leaf records (a 420K-key domain
    # expecting few groups) and the dense leaf output (SMALLINT key)
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    st = run_both(session, t, Request(["s", "k2"], ALL_AGGS[:6], [SortSpec("c", True)], 20, est_groups=10), small)
    assert st.global_mode == 3
    st = run_both(session, t, Request(["s"], ALL_AGGS[:6]), small)
    assert st.global_mode == 3


def test_noalloc_limits(session):
    """An execution whose buffers exceed the scratch array, and an output
    buffer too small, fail with their status and a message; a zero-row scan
    returns no rows."""
    rng = np.random.default_rng(9)
    t = random_table(rng, 20_000, rg_rows=1000, nulls=False)
    plan, cols = build_plan(session, t, Request(["k1"], ALL_AGGS[:4]))
    session.stage_table(t, cols)
    try:
        with pytest.raises(gpu.GpuError, match="INVALID.*output bytes"):
            plan.execute_noalloc(16)
        assert len(rows(t, Request(["k1"], ALL_AGGS[:4]), plan.execute_noalloc(CAP))) == 100
    finally:
        plan.close()
    none = Request(["k1"], ALL_AGGS[:4], where_filter=Predicate("i", "gt", 2 ** 31))  # every row group pruned
    plan, cols = build_plan(session, t, none)
    session.stage_table(t, cols)
    try:
        assert [len(c.values) for c in plan.execute_noalloc(CAP)] == [0] * 5
    finally:
        plan.close()
    # 2^26 dense records of 16+ bytes: over 64 MiB
    cols_t = dict(t.columns)
    cols_t["t"] = np.arange(t.n_rows, dtype=np.int64) * 3_355  # span ~ 6.7e7
    from thindb_amd.exec import Table
    wide = Table(t.schema, cols_t, t.validity, rg_rows=1000)
    plan, cols = build_plan(session, wide, Request(["t"], [AggSpec("count", None, "c")], est_groups=1 << 26))
    session.stage_table(wide, cols)
    try:
        with pytest.raises(gpu.Unsupported, match="scratch"):
            plan.execute_noalloc(CAP)
    finally:
        plan.close()
