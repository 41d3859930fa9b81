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


def float_domain(rng):
    """+0.0, -0.0, four NaNs (quiet and signalling payloads, both signs) and
    45 ordinary values: 47 distinct values once canonicalised."""
    nans = np.array([0x7FF8000000000000, 0x7FF8000000000123, 0xFFF8000000000000, 0x7FF0000000000001],
                    dtype=np.uint64).view(np.float64)
    return np.concatenate([[1.5, -2.25, 0.0, -0.0, 1e300, -1e-300, 3.0], nans, rng.uniform(-100, 100, 40)])


def float_table(seed=6, n=40_000):
    """k1 INT keys; d DOUBLE (nullable) and f FLOAT over small domains that
    hold +0.0 and -0.0 and NaNs of several payloads and signs; w BIGINT
    (nullable) over values spread across the whole int64 range (its span
    does not fit a 63-bit composite); g BIGINT, 500 keys sparse in +-2^40."""
    rng = np.random.default_rng(seed)
    dvals = float_domain(rng)
    fnans = np.array([0x7FC00000, 0x7FC00123, 0xFFC00000, 0x7F800001], dtype=np.uint32).view(np.float32)
    fvals = np.concatenate([np.array([1.5, -2.25, 0.0, -0.0, 3.0e38, 1e-40], dtype=np.float32), fnans,
                            rng.normal(0, 10, 30).astype(np.float32)])
    wvals = np.concatenate([[-2 ** 63, 2 ** 63 - 1, 0, -1], rng.integers(-2 ** 62, 2 ** 62, 200, dtype=np.int64)])
    gvals = rng.integers(-2 ** 40, 2 ** 40, 500, dtype=np.int64)
    cols = {"k1": rng.integers(-20, 20, n).astype(np.int32), "d": dvals[rng.integers(0, len(dvals), n)],
            "f": fvals[rng.integers(0, len(fvals), n)], "w": wvals[rng.integers(0, len(wvals), n)].astype(np.int64),
            "g": gvals[rng.integers(0, len(gvals), n)]}
    schema = [Column("k1", Type.INT), Column("d", Type.DOUBLE, True), Column("f", Type.FLOAT),
              Column("w", Type.BIGINT, True), Column("g", Type.BIGINT)]
    return Table(schema, cols, {"d": rng.random(n) > 0.15, "w": rng.random(n) > 0.1}, rg_rows=997)


FLOAT_AGGS = [AggSpec("count", None, "c"), AggSpec("count_distinct", "d", "dd"), AggSpec("count_distinct", "f", "df"),
              AggSpec("count_distinct", "w", "dw")]


def test_oracle_count_distinct_floats_and_wide_values(oracle_lib):
    """Float values count one +0.0 and one NaN (every NaN payload and sign is
    one value), a FLOAT widens to f64; BIGINT values across the whole range
    count exactly."""
    t = float_table()
    for keys, wf in ((["k1"], None), (["k1"], Predicate("w", "gt", 0)), (["g"], None), (["k1", "g"], None)):
        req = Request(keys, FLOAT_AGGS, where_filter=wf)
        want = brute_force(t, req)
        got = rows_of(oracle_lib.run(t, req, dop=3))
        assert len(got) == len(want)
        for r in got:
            assert list(r[len(keys):]) == want[tuple(r[:len(keys)])], (keys, r)
    # the closed form over one group: 47 distinct values in the domain
    d = float_domain(np.random.default_rng(0))
    one = Table([Column("k", Type.INT), Column("d", Type.DOUBLE)], {"k": np.zeros(len(d), np.int32), "d": d})
    assert rows_of(oracle_lib.run(one, Request(["k"], [AggSpec("count_distinct", "d", "x")]), dop=1)) == [(0, 47)]


def test_count_distinct_schema_and_declines():
    from thindb_amd.exec import Declined, _having_postfix  # noqa: F401
    t = _table()
    sch = output_schema(t, Request(["k1"], AGGS))
    assert [(c.name, c.type, c.nullable) for c in sch[1:3]] == [("c", Type.BIGINT, False), ("dk2", Type.BIGINT, False)]
