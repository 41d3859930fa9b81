"""Nullable GROUP BY keys on the CPU: the oracle against brute force.

Semantics (src/exec/v2_lowcard_group.zig:44-49, :886-966, :2219-2224;
src/exec/v2_shape_group_topn.zig:1475-1479): a NULL key value forms its own
group, apart from the value 0; the key comes out NULL; ORDER BY the key puts
NULL first ascending and last descending; HAVING over a NULL key fails its
leaf and IS NULL holds.
"""
import numpy as np

from helpers import brute_force, rows_of
from thindb_amd.exec import (AggSpec, Column, IsNull, Not, Or, Predicate, Request, SortSpec, Table,
                             output_schema)
from thindb_amd.gpu import Type

AGGS = [AggSpec("count", None, "c"), AggSpec("sum", "v", "s"), AggSpec("min", "v", "mn"),
        AggSpec("count_distinct", "w", "dw"), AggSpec("avg", "f", "a")]


def null_key_table(seed=7, n=40_000, rg_rows=997):
    rng = np.random.default_rng(seed)
    schema = [Column("k1", Type.INT, True), Column("k2", Type.SMALLINT, True), Column("k3", Type.BIGINT),
              Column("v", Type.BIGINT, True), Column("w", Type.INT), Column("f", Type.DOUBLE)]
    cols = {"k1": rng.integers(-5, 6, n).astype(np.int32), "k2": rng.integers(0, 4, n).astype(np.int16),
            "k3": rng.integers(10, 14, n).astype(np.int64), "v": rng.integers(-10 ** 6, 10 ** 6, n),
            "w": rng.integers(0, 40, n).astype(np.int32), "f": rng.random(n)}
    validity = {"k1": rng.random(n) > 0.2, "k2": rng.random(n) > 0.3, "v": rng.random(n) > 0.1}
    # NULL slots hold values outside the valid range, which must not matter
    cols["k1"][~validity["k1"]] = 1_000_000
    cols["k2"][~validity["k2"]] = -77
    return Table(schema, cols, validity, rg_rows=rg_rows)


def _sorted_keys(rows, nk):
    return sorted((tuple(r[:nk]) for r in rows), key=lambda k: tuple((x is not None, x or 0) for x in k))


def test_oracle_nullable_keys_match_brute_force(oracle_lib):
    t = null_key_table()
    for keys in (["k1"], ["k2", "k1"], ["k1", "k3"], ["k2"]):
        req = Request(keys, AGGS)
        want = brute_force(t, req)
        got = rows_of(oracle_lib.run(t, req, dop=3))
        assert len(got) == len(want)
        assert any(None in k for k in want)            # NULL groups exist
        for r in got:
            w = want[tuple(r[:len(keys)])]
            assert list(r[len(keys):len(keys) + 4]) == w[:4], (keys, r)
            assert abs(r[-1] - w[-1]) <= 1e-12 * abs(w[-1])
    got = {r[0] for r in rows_of(oracle_lib.run(t, Request(["k1"], AGGS), dop=1))}
    assert None in got and 0 in got and 1_000_000 not in got


def test_oracle_nullable_key_order_and_having(oracle_lib):
    t = null_key_table(seed=8)
    asc = rows_of(oracle_lib.run(t, Request(["k1"], AGGS, [SortSpec("k1")]), dop=2))
    assert asc[0][0] is None and [r[0] for r in asc[1:]] == sorted(r[0] for r in asc[1:])
    desc = rows_of(oracle_lib.run(t, Request(["k1"], AGGS, [SortSpec("k1", True)], limit=20), dop=2))
    assert desc[-1][0] is None and desc[0][0] == 5
    # two keys: NULL first within each k2
    two = rows_of(oracle_lib.run(t, Request(["k2", "k1"], AGGS, [SortSpec("k2"), SortSpec("k1")]), dop=2))
    assert _sorted_keys(two, 2) == [tuple(r[:2]) for r in two]
    n_null = sum(1 for r in two if r[1] is None)
    h = rows_of(oracle_lib.run(t, Request(["k2", "k1"], AGGS, having_filter=IsNull("k1")), dop=2))
    assert len(h) == n_null and all(r[1] is None for r in h)
    gt = rows_of(oracle_lib.run(t, Request(["k2", "k1"], AGGS, having_filter=Predicate("k1", "gt", 0)), dop=2))
    ngt = rows_of(oracle_lib.run(t, Request(["k2", "k1"], AGGS, having_filter=Not(Predicate("k1", "gt", 0))), dop=2))
    assert all(r[1] is not None and r[1] > 0 for r in gt)
    assert len(gt) + len(ngt) == len(two) and sum(r[1] is None for r in ngt) == n_null
    either = rows_of(oracle_lib.run(t, Request(["k2", "k1"], AGGS, having_filter=Or((IsNull("k1"), IsNull("k2")))),
                                    dop=2))
    assert len(either) == sum(1 for r in two if r[0] is None or r[1] is None)


def test_nullable_key_schema():
    t = null_key_table()
    sch = output_schema(t, Request(["k1", "k3"], AGGS))
    assert [c.nullable for c in sch[:2]] == [True, False]
