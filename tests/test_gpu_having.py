"""HAVING on the GPU, through the C-ABI, against the oracle: every group
strategy (direct LDS table, HBM hash table, radix partitions with dense and
compacted leaf output, the batched launch), with ORDER BY / OFFSET / LIMIT
after the filter, NULL aggregates, and malformed node arrays rejected."""
from decimal import Decimal

import numpy as np
import pytest

from helpers import ALL_AGGS, assert_same_rows, rows_of
from test_gpu_parity import REL, check_vs_oracle, float_cols_of, run_gpu
from test_having import _cases, _table
from thindb_amd import gen, gpu
from thindb_amd.exec import (And, AggSpec, Column, GpuBatchRun, IsNull, Not, Or, Predicate, Request, SortSpec,
                             Table)
from thindb_amd.gpu import HavingKind, PredOp, Type

pytestmark = pytest.mark.gpu


def test_having_cases_vs_oracle(session, oracle_lib):
    """Every HAVING tree of the CPU suite over one and two keys, unordered
    and under ORDER BY ... LIMIT."""
    t = _table()
    for having in _cases(t):
        print(f"[having] {having}", flush=True)
        check_vs_oracle(session, oracle_lib, t, Request(["k1"], ALL_AGGS, having_filter=having))
        check_vs_oracle(session, oracle_lib, t, Request(["k1", "k2"], ALL_AGGS, [SortSpec("sb", True)], 9, 2,
                                                        having_filter=having))


def test_having_hash_table(session, oracle_lib):
    rng = np.random.default_rng(5)
    n = 300_000
    keys = rng.integers(-2 ** 40, 2 ** 40, 4000)
    t = Table([Column("k", Type.BIGINT), Column("v", Type.INT), Column("d", Type.DOUBLE)],
              {"k": keys[rng.integers(0, 4000, n)], "v": rng.integers(-100, 100, n).astype(np.int32),
               "d": rng.random(n)})
    aggs = [AggSpec("count", None, "c"), AggSpec("sum", "v", "s"), AggSpec("avg", "d", "a"), AggSpec("min", "v", "m")]
    h = And((Predicate("c", "gte", 75), Or((Predicate("s", "lt", -50), Predicate("k", "gt", 0)))))
    got, want, st = check_vs_oracle(session, oracle_lib, t, Request(["k"], aggs, having_filter=h))
    assert st.global_mode == 2 and 0 < len(got) < 4000
    check_vs_oracle(session, oracle_lib, t, Request(["k"], aggs, [SortSpec("a", True)], 25, having_filter=h))


@pytest.mark.parametrize("est", [0, 100_000_000])
def test_having_radix(session, oracle_lib, monkeypatch, est):
    """The radix path with compacted leaf output (default) and dense output
    (a large est_groups); COUNT(*)-ordered top-N goes through the general
    sort when HAVING is present."""
    t = gen.table("highcard", 1_500_000, seed=12)
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    aggs = [AggSpec("count", None, "c"), AggSpec("sum", "v", "s"), AggSpec("max", "w", "m")]
    h = Or((Predicate("c", "gte", 2), Predicate("s", "lt", 0)))
    got, want, st = check_vs_oracle(session, oracle_lib, t, Request(["user_id", "bucket"], aggs, est_groups=est,
                                                                    having_filter=h), dop=8)
    assert st.global_mode == 3 and st.n_groups == len(want) > 0
    check_vs_oracle(session, oracle_lib, t, Request(["user_id", "bucket"], aggs, [SortSpec("c", True)], 10,
                                                    est_groups=est, having_filter=Not(h)), dop=8)
    check_vs_oracle(session, oracle_lib, t, Request(["user_id", "bucket"], aggs, [SortSpec("c", True)], 100,
                                                    est_groups=est, having_filter=h), dop=8)


def test_having_in_radix_leaves(session, oracle_lib, monkeypatch):
    """Every HAVING tree evaluated by the radix leaf pass on its LDS table
    (compacted leaf output: a small est_groups), with and without ORDER BY."""
    t = _table()
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    for having in _cases(t):
        print(f"[having/leaf] {having}", flush=True)
        _, _, st = check_vs_oracle(session, oracle_lib, t, Request(["k1", "k2"], ALL_AGGS, est_groups=10,
                                                                   having_filter=having))
        assert st.global_mode == 3
        # COUNT ties broken by the keys, so the OFFSET cuts a total order
        check_vs_oracle(session, oracle_lib, t, Request(["k2", "k1"], ALL_AGGS,
                                                        [SortSpec("c", True), SortSpec("k2"), SortSpec("k1")], 7, 1,
                                                        est_groups=10, having_filter=having))


def test_having_filtered_workload(session, oracle_lib):
    """The filtered scan workload (WHERE on ts) with HAVING on the DECIMAL
    SUM and the AVG, ordered by the SUM."""
    t = gen.table("filtered", 2_000_000, seed=2)
    where = And((Predicate("ts", "gte", gen.FILTER_LO), Predicate("ts", "lt", gen.FILTER_HI)))
    aggs = [AggSpec("count", None, "c"), AggSpec("sum", "amount", "s"), AggSpec("avg", "price", "a")]
    every, _ = run_gpu(session, t, Request(["region"], aggs, where_filter=where))
    sums = sorted(r[2] for r in rows_of(every))
    cut = Decimal((sums[len(sums) // 2] + sums[len(sums) // 2 + 1]) // 2).scaleb(-t.column("amount").scale)
    h = Predicate("s", "gt", cut)
    got, want, _ = check_vs_oracle(session, oracle_lib, t, Request(["region"], aggs, [SortSpec("s", True)], 100,
                                                                   where_filter=where, having_filter=h), dop=8)
    assert len(got) == min(100, len(sums) - len(sums) // 2 - 1)


def test_having_empty_and_null(session, oracle_lib):
    """A HAVING no group passes; one over the NULL aggregates of groups with
    no non-null input; an empty scan."""
    t = _table()
    dead = [-50, -49, -48, -47]
    got, _, _ = check_vs_oracle(session, oracle_lib, t, Request(["k1"], ALL_AGGS,
                                                                having_filter=Predicate("c", "lt", 0)))
    assert got == []
    got, _, _ = check_vs_oracle(session, oracle_lib, t, Request(["k1"], ALL_AGGS, [SortSpec("k1")],
                                                                having_filter=IsNull("sb")))
    assert [r[0] for r in got] == dead
    got, _, _ = check_vs_oracle(session, oracle_lib, t, Request(["k1"], ALL_AGGS, [SortSpec("k1")],
                                                                having_filter=Not(Predicate("mins", "gt", -10 ** 6))))
    assert [r[0] for r in got] == dead
    got, _, _ = check_vs_oracle(session, oracle_lib, t, Request(["k1"], ALL_AGGS,
                                                                where_filter=Predicate("i", "gt", 2 ** 31),
                                                                having_filter=Predicate("c", "gt", 0)))
    assert got == []


def test_having_batched(session, oracle_lib):
    """Requests with HAVING inside one batched launch return what each
    returns alone and what the oracle returns."""
    t = _table()
    hs = _cases(t)[:8]
    reqs = [Request(["k1"] if i % 2 else ["k1", "k2"], ALL_AGGS, having_filter=h) for i, h in enumerate(hs)]
    run = GpuBatchRun(session, [(t, r) for r in reqs])
    try:
        out = run.execute()
        for r, b in zip(reqs, out):
            want = rows_of(oracle_lib.run(t, r, 4))
            assert_same_rows(rows_of(b), want, float_cols_of(t, r), REL, len(r.group_cols))
    finally:
        run.close()


def test_having_rejects_malformed_nodes(session):
    """The library rejects node arrays that do not form one tree."""
    t = _table()
    session.stage_table(t, ["k1"])
    keys = [(t.column_index("k1"), Type.INT)]
    aggs = [(gpu.AggOp.COUNT_STAR, -1, Type.BIGINT, 0, False)]
    leaf = (HavingKind.CMP, 1, PredOp.GT, 3)
    bad = [
        [(HavingKind.NOT,)],                              # NOT of nothing
        [leaf, leaf],                                     # two entries left
        [leaf, (HavingKind.AND, 2)],                      # arity beyond the stack
        [(HavingKind.CMP, 2, PredOp.GT, 3)],              # no output column 2
        [(HavingKind.CMP, 1, 9, 3)],                      # unknown op
        [(HavingKind.IS_NULL, 5)],
    ]
    for nodes in bad:
        with pytest.raises(gpu.InvalidArgument):
            gpu.Plan(session.ctx, t.uid, keys, aggs, having=nodes)
    with pytest.raises(gpu.Unsupported):
        gpu.Plan(session.ctx, t.uid, keys, aggs, having=[leaf] * 33 + [(HavingKind.AND, 33)])
    p = gpu.Plan(session.ctx, t.uid, keys, aggs, having=[leaf, (HavingKind.AND, 1), (HavingKind.OR, 0),
                                                         (HavingKind.OR, 2)])
    try:
        got = p.execute()
        assert len(got[0].values) == 100 and (np.asarray(got[1].values) > 3).all()
    finally:
        p.close()
