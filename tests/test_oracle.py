"""Pins the CPU oracle before it is trusted as the checker.

* the reference's own known-answer tests (tests/golden/*.json);
* an independent pure-Python brute force over random typed tables with
  NULLs and filters (exact integers, math.fsum for floats);
* invariants: DOP-1 = DOP-N, sum of COUNT = surviving rows;
* planted mistakes in a copy of the oracle's plan must be caught.
"""
import math

import numpy as np
import pytest

from helpers import (ALL_AGGS, assert_rows_in_groups, brute_force, decimal_table, float_table, golden,
                     heavy_table, random_table, rows_of, wrap_table)
from thindb_amd.exec import And, AggSpec, Predicate, Request, SortSpec


def test_wrap_sum_known_answers(oracle_lib):
    g = golden("wrap_sum")
    t = wrap_table()
    for q in g["queries"]:
        having = Predicate(*q["having"]) if "having" in q else None
        req = Request(q["keys"], [AggSpec("sum", "v", "s")], [SortSpec(c, d) for c, d in q["order"]], q["limit"],
                      having_filter=having)
        for dop in (1, 4):
            got = rows_of(oracle_lib.run(t, req, dop))
            assert [list(r) for r in got] == q["expect"], (q["sql"], dop)
    res = oracle_lib.run(t, Request(["hk"], [AggSpec("sum", "v", "s")]), 1)
    rows = rows_of(res)
    assert len(rows) == g["all_hk_groups"]
    want = {1: g["WRAP_G1"], 2: g["WRAP_G2"], 3: g["WRAP_G3"]}
    for k, s in rows:
        assert s == want.get(k, 0)


def test_decimal_sum_avg_known_answers(oracle_lib):
    g = golden("decimal_group")
    t = decimal_table()
    res = oracle_lib.run(t, Request(["grp"], [AggSpec("sum", "v", "sv"), AggSpec("avg", "v", "av")],
                                    [SortSpec("grp")]), 1)
    assert res.columns[1].type == 11 and res.columns[1].scale == 4  # DECIMAL128, scale kept
    rows = rows_of(res)
    assert [[r[0], r[1]] for r in rows] == g["sum"]
    for (k, want, tol), r in zip(g["avg_approx"], rows):
        assert abs(r[2] - want) <= tol
    for (k, exact), r in zip(g["avg_exact"], rows):
        assert r[2] == exact


def test_float_aggregates_known_answers(oracle_lib):
    g = golden("float_group")
    t = float_table()
    aggs = [AggSpec("sum", "sd", "s"), AggSpec("min", "sd", "mn"), AggSpec("max", "sd", "mx"),
            AggSpec("count", None, "c"), AggSpec("avg", "sd", "a"), AggSpec("sum", "sf", "sf"),
            AggSpec("min", "sf", "mnf"), AggSpec("max", "sf", "mxf")]
    rows = rows_of(oracle_lib.run(t, Request(["k"], aggs, [SortSpec("k")]), 1))
    for r in rows:
        e = g["expect"][str(r[0])]
        assert (r[1], r[2], r[3], r[4]) == (e["sum"], e["min"], e["max"], e["count"])
        assert r[5] == e["avg"]
        assert (r[6], r[7], r[8]) == (e["sum"], e["min"], e["max"])


def test_heavy_groups_count_topk(oracle_lib):
    g = golden("heavy_groups")
    t = heavy_table()
    req = Request(["k"], [AggSpec("count", None, "c"), AggSpec("sum", "v", "s")], [SortSpec("c", True)], 10)
    for dop in (1, 4):
        rows = rows_of(oracle_lib.run(t, req, dop))
        assert [list(r) for r in rows] == g["top10"]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_brute_force_random_tables(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    t = random_table(rng, 3000, rg_rows=256)
    where = And((Predicate("i", "gt", -2 ** 30), Predicate("t", "lt", 7 * 10 ** 11), Predicate("s", "neq", 0)))
    for keys, wf in [(["k1"], None), (["k1", "k2"], where), (["k2"], where)]:
        req = Request(keys, ALL_AGGS, where_filter=wf)
        want = brute_force(t, req)
        res = oracle_lib.run(t, req, 1)
        rows = rows_of(res)
        assert len(rows) == len(want) == res.n_groups
        nk = len(keys)
        float_cols = [nk + j for j, a in enumerate(ALL_AGGS) if a.col in ("d", "f") and a.func in ("sum", "avg")]
        assert_rows_in_groups(rows, want, nk, float_cols, rel=1e-12)
        # compensated sums track fsum to the last bits
        for j in float_cols:
            comp = res.columns[j].compensated
            for r, row in enumerate(rows):
                w = want[tuple(row[:nk])][j - nk]
                if w is not None and w != 0:
                    assert abs(comp[r] - w) <= 4e-16 * abs(w) + 1e-300


def test_dop_invariance_and_count_total(oracle_lib):
    rng = np.random.default_rng(7)
    t = random_table(rng, 20000, rg_rows=512, nulls=True)
    req = Request(["k1", "k2"], [AggSpec("count", None, "c"), AggSpec("sum", "b", "s"), AggSpec("min", "i", "m"),
                                 AggSpec("max", "s", "x")], [SortSpec("k1"), SortSpec("k2")])
    a = rows_of(oracle_lib.run(t, req, 1))
    b = rows_of(oracle_lib.run(t, req, 4))
    assert a == b
    assert sum(r[2] for r in a) == t.n_rows


def test_order_offset_limit_and_ties(oracle_lib):
    rng = np.random.default_rng(3)
    t = random_table(rng, 5000, rg_rows=500, nulls=False)
    req = Request(["k1"], [AggSpec("count", None, "c"), AggSpec("sum", "m", "sm")],
                  [SortSpec("c", True), SortSpec("k1")], limit=7, offset=3)
    full = rows_of(oracle_lib.run(t, Request(["k1"], req.aggs, req.order_specs), 1))
    page = rows_of(oracle_lib.run(t, req, 1))
    assert page == full[3:10]
    # every ranked row is in order (desc count, then asc key)
    assert full == sorted(full, key=lambda r: (-r[1], r[0]))


def test_declines_like_the_reference(oracle_lib):
    rng = np.random.default_rng(0)
    t = random_table(rng, 100)
    with pytest.raises(oracle_lib.Unsupported):
        oracle_lib.run(t, Request(["k1"], [AggSpec("sum", "t", "x")]), 1)  # temporal SUM
    with pytest.raises(oracle_lib.Unsupported):
        oracle_lib.run(t, Request(["d"], [AggSpec("count", None, "c")]), 1)  # float key


def test_planted_mistakes_are_caught(oracle_lib):
    """A dropped row, a wrong sign and a lost carry each break a check the
    suite relies on (run against the brute force, on the CPU)."""
    rng = np.random.default_rng(11)
    t = random_table(rng, 2000, rg_rows=256, nulls=False)
    req = Request(["k2"], [AggSpec("count", None, "c"), AggSpec("sum", "b", "s")])
    want = brute_force(t, req)
    rows = rows_of(oracle_lib.run(t, req, 1))
    assert_rows_in_groups(rows, want, 1)
    bad_drop = [(k, c - 1, s) if k == rows[0][0] else (k, c, s) for k, c, s in rows]
    bad_sign = [(k, c, -s) if k == rows[1][0] else (k, c, s) for k, c, s in rows]
    bad_carry = [(k, c, ((s + (1 << 62)) & ((1 << 64) - 1)) - (1 << 64) * (((s + (1 << 62)) >> 63) & 1))
                 if k == rows[2][0] else (k, c, s) for k, c, s in rows]
    for bad in (bad_drop, bad_sign, bad_carry):
        with pytest.raises(AssertionError):
            assert_rows_in_groups(bad, want, 1)


@pytest.mark.parametrize("seed", range(4))
def test_radix_arrangement_equals_the_oracle(oracle_lib, seed):
    """oracle_group_by_radix (the silo's route / bucket / part arrangement, the
    CPU baseline for high-cardinality queries) returns the oracle's rows:
    random typed tables with NULLs and filters, every aggregate kind, one to
    two keys, ORDER BY with OFFSET/LIMIT (the top-k selection) and HAVING;
    and a high-cardinality table whose buckets split into parts."""
    from helpers import float_cols_of, order_cols_of, unsliced
    rng = np.random.default_rng(300 + seed)
    t = random_table(rng, 30_000 + 5_000 * seed, rg_rows=1_000 + 137 * seed)
    aggs = [a for a in ALL_AGGS if a.func != "count_distinct"]
    reqs = [Request(["k1"], aggs), Request(["k1", "k2"], aggs, [SortSpec(aggs[1].as_, True)], 25, 3),
            Request(["k2", "k1"], aggs[:5], where_filter=Predicate("i", "gt", 0)),
            Request(["k1", "k2"], aggs[:4], [SortSpec("k1", True)], 7,
                    having_filter=Predicate(aggs[0].as_, "gte", 2))]
    for req in reqs:
        want = rows_of(oracle_lib.run(t, req, 4))
        got = rows_of(oracle_lib.run(t, req, 4, radix=True))
        names = [k for k in req.group_cols] + [a.as_ for a in req.aggs]
        full = rows_of(oracle_lib.run(t, unsliced(req), 4)) if req.limit or req.offset else None
        from helpers import assert_same_rows
        assert_same_rows(got, want, float_cols_of(t, req), 1e-12, len(req.group_cols), order_cols_of(names, req),
                         full)
    # many groups: every bucket splits into parts
    n = 400_000
    from thindb_amd.exec import Column, Table
    from thindb_amd.gpu import Type
    hc = Table([Column("u", Type.BIGINT), Column("b", Type.TINYINT), Column("v", Type.INT)],
               {"u": rng.integers(0, 150_000, n), "b": rng.integers(0, 4, n).astype(np.int8),
                "v": rng.integers(-1000, 1000, n).astype(np.int32)})
    req = Request(["u", "b"], [AggSpec("count", None, "c"), AggSpec("sum", "v", "s"), AggSpec("max", "v", "m")],
                  [SortSpec("c", True)], 10)
    assert rows_of(oracle_lib.run(hc, req, 8, radix=True)) == rows_of(oracle_lib.run(hc, req, 8))
    r = Request(["u", "b"], req.aggs)
    a = sorted(rows_of(oracle_lib.run(hc, r, 8, radix=True)))
    assert a == sorted(rows_of(oracle_lib.run(hc, r, 1)))
