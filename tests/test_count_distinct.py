"""COUNT(DISTINCT col) on the CPU: the oracle's per-group distinct counts
pinned against brute force (len(set(...)) of each group's non-NULL values),
with NULLs, filters, several keys, ORDER BY / LIMIT and HAVING over the
count, and the operator's declines.

Semantics (src/exec/v2_lowcard_group.zig:363-387, :1348-1418): distinct
(group, value) pairs of the filtered rows, NULL values skipped; the result is
BIGINT and never NULL (0 when a group's values are all NULL).
"""
import numpy as np
import pytest

from helpers import brute_force, random_table, rows_of
from thindb_amd.exec import (AggSpec, And, Column, Predicate, Request, SortSpec, Table, output_schema)
from thindb_amd.gpu import Type

AGGS = [AggSpec("count", None, "c"), AggSpec("count_distinct", "k2", "dk2"), AggSpec("count_distinct", "s", "ds"),
        AggSpec("sum", "i", "si"), AggSpec("count_distinct", "m", "dm"), AggSpec("count_distinct", "t", "dt")]


def _table(seed=3, n=30_000):
    rng = np.random.default_rng(seed)
    t = random_table(rng, n, rg_rows=1001, nulls=True)
    # small value domains so groups repeat values; the keys k1 < -45 hold
    # only NULL s (their distinct count is 0)
    t.columns["s"] = rng.integers(-40, 40, n).astype(np.int16)
    t.columns["m"] = rng.integers(0, 300, n).astype(np.int64)
    t.columns["t"] = (rng.integers(0, 50, n) * 10 ** 9).astype(np.int64)
    t.validity["s"] = t.validity["s"] & ~(t.columns["k1"] < -45)
    t._packed.clear()
    return t


def test_oracle_count_distinct_matches_brute_force(oracle_lib):
    t = _table()
    where = And((Predicate("i", "gt", -2 ** 30), Predicate("d", "gte", -500.0)))
    for keys, wf in ((["k1"], None), (["k1", "k2"], None), (["k2"], where), (["k1"], where)):
        req = Request(keys, AGGS, where_filter=wf)
        want = brute_force(t, req)
        got = rows_of(oracle_lib.run(t, req, dop=3))
        assert len(got) == len(want)
        for r in got:
            assert list(r[len(keys):]) == want[tuple(r[:len(keys)])], (keys, r)
    # the all-NULL groups count 0, never NULL
    got = {r[0]: r for r in rows_of(oracle_lib.run(t, Request(["k1"], AGGS), dop=1))}
    assert all(got[k][3] == 0 for k in range(-50, -45))
    assert all(got[k][3] > 0 for k in range(-45, 50))


def test_oracle_count_distinct_order_and_having(oracle_lib):
    """ORDER BY the distinct count (ties by the key) and HAVING on it equal
    the brute-force groups ordered / filtered."""
    t = _table(seed=4)
    req_all = Request(["k1", "k2"], AGGS)
    bf = brute_force(t, req_all)
    order = sorted(bf.items(), key=lambda kv: (-kv[1][2], kv[0]))
    req = Request(["k1", "k2"], AGGS, [SortSpec("ds", True), SortSpec("k1"), SortSpec("k2")], limit=25, offset=5)
    got = rows_of(oracle_lib.run(t, req, dop=2))
    assert [tuple(r[:2]) for r in got] == [k for k, _ in order[5:30]]
    h = Predicate("dm", "gte", 40)
    got = rows_of(oracle_lib.run(t, Request(["k1", "k2"], AGGS, having_filter=h), dop=2))
    assert sorted(tuple(r[:2]) for r in got) == sorted(k for k, v in bf.items() if v[4] >= 40)


def test_oracle_count_distinct_invariants(oracle_lib):
    """COUNT(DISTINCT k) over a key column is 1 per group; over a column
    constant per group, 1; distinct counts never exceed COUNT(col)."""
    t = _table(seed=5)
    got = rows_of(oracle_lib.run(t, Request(["k1"], [AggSpec("count_distinct", "k1", "d"),
                                                     AggSpec("count", "s", "cs"),
                                                     AggSpec("count_distinct", "s", "ds")]), dop=2))
    assert all(r[1] == 1 for r in got)
    assert all(r[3] <= r[2] for r in got)
    n = 5000
    k = np.arange(n, dtype=np.int32) % 17
    t2 = Table([Column("k", Type.INT), Column("v", Type.BIGINT)], {"k": k, "v": (k.astype(np.int64) * 7 + 3)})
    got = rows_of(oracle_lib.run(t2, Request(["k"], [AggSpec("count_distinct", "v", "d")]), dop=1))
    assert sorted(got) == [(i, 1) for i in range(17)]


def test_count_distinct_schema_and_declines():
    from thindb_amd.exec import Declined, _having_postfix  # noqa: F401
    t = _table()
    sch = output_schema(t, Request(["k1"], AGGS))
    assert [(c.name, c.type, c.nullable) for c in sch[1:3]] == [("c", Type.BIGINT, False), ("dk2", Type.BIGINT, False)]
