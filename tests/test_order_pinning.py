"""Pins the oracle's ORDER BY comparator (CPU only).

* The reference's ranked-emit pages (tests/integration/v2_group_topn_test.zig:
  1408-1460: 140,009 groups, MIN / SUM with HAVING / COUNT / AVG / SUM
  orders with LIMIT and OFFSET, ties broken by the group key) and its
  nullable-aggregate pages (:1283-1400), as fixtures written by
  tests/golden/make_golden.py from the tests' own rules.
* An independent brute force (helpers.brute_order: the groups from the
  definitions, sorted by a comparator written from the silo's
  rowOrderBefore) over a table built to separate comparators: keys of both
  signs and a nullable key (the packed-key tie-break), DECIMAL sums beyond 64
  bits, all-NULL aggregates, integer AVGs, DOUBLE sums holding NaN and
  signed zeros, COUNT ties, multi-item orders.
* Planted mistakes in the oracle's comparator (oracle.planted) each fail
  one of these checks.
"""
import hashlib
import json

import pytest

from helpers import (brute_order, golden, nullable_agg_table, order_table, ranked_table, rows_of, same_values,
                     wide_key_table)
from thindb_amd.exec import AggSpec, Predicate, Request, SortSpec

def ranked_request(case):
    """One case of ranked_pages.json; a HAVING COUNT(*) > 1 adds the count as
    a third output (the page compares the first two)."""
    func = case["func"]
    name = {"min": "m", "sum": "s", "count": "c", "avg": "a"}[func]
    aggs = [AggSpec("count", None, "c") if func == "count" else AggSpec(func, "v", name)]
    having = None
    if case["having_count_gt_1"]:
        aggs.append(AggSpec("count", None, "n"))
        having = Predicate("n", "gt", 1)
    sql = case["sql"]
    order_sql = sql[sql.index("ORDER BY ") + 9:sql.index(" LIMIT")]
    order = []
    for item in order_sql.split(", "):
        col, *d = item.split(" ")
        order.append(SortSpec(col, bool(d and d[0] == "DESC")))
    return Request(["g"], aggs, order, case["limit"], case["offset"], having_filter=having)


RANKED_CASES = golden("ranked_pages")["cases"]


@pytest.fixture(scope="module")
def ranked():
    return ranked_table()


@pytest.mark.parametrize("i", range(len(RANKED_CASES)))
def test_ranked_emit_pages(oracle_lib, ranked, i):
    case = RANKED_CASES[i]
    req = ranked_request(case)
    for dop in (1, 4):
        rows = rows_of(oracle_lib.run(ranked, req, dop))
        assert [[r[0], r[1]] for r in rows] == case["page"], (case["sql"], dop)


NULLABLE_SQL = {
    "int": [("sum", "id", "si"), ("count", None, "c"), ("sum", "v", "sv"), ("avg", "v", "av"), ("min", "v", "mnv"),
            ("max", "v", "mxv"), ("sum", "b", "sb")],
    "float": [("sum", "id", "si"), ("avg", "b", "ab"), ("sum", "d", "sd"), ("min", "d", "mnd"), ("max", "d", "mxd")],
}


def nullable_request(kind, limit=0, offset=0):
    aggs = [AggSpec(f, c, a) for f, c, a in NULLABLE_SQL[kind]]
    return Request(["g"], aggs, [SortSpec("si", True)], limit, offset)


@pytest.mark.parametrize("kind", ["int", "float"])
def test_nullable_aggregate_pages(oracle_lib, kind):
    g = golden("nullable_aggs")
    t = nullable_agg_table()
    for dop in (1, 4):
        page = rows_of(oracle_lib.run(t, nullable_request(kind, 25, 7), dop))
        assert [list(r) for r in page] == g[kind]["page_25_7"]
        every = rows_of(oracle_lib.run(t, nullable_request(kind), dop))
        assert len(every) == g["group_count"]
        assert hashlib.sha256(json.dumps([list(r) for r in every]).encode()).hexdigest() == g[kind]["digest_all"]


# ---- brute-force comparator ----------------------------------------------------------------

ORDER_AGGS = [AggSpec("count", None, "c"), AggSpec("sum", "m", "sm"), AggSpec("avg", "m", "am"),
              AggSpec("min", "s", "mins"), AggSpec("sum", "s", "ss"), AggSpec("avg", "i", "ai"),
              AggSpec("sum", "d", "sd"), AggSpec("avg", "d", "ad"), AggSpec("max", "e", "mxe"),
              AggSpec("min", "e", "mne"), AggSpec("sum", "i", "si")]

ORDER_CASES = [
    (["k1", "k2"], [SortSpec("sm", True)], 10, 0),             # DECIMAL128 SUM beyond 64 bits
    (["k1"], [SortSpec("ai")], 7, 0),                          # integer AVG: same integer part, other fraction
    (["k1", "k2"], [SortSpec("mins")], 15, 0),                 # all-NULL MIN first ascending
    (["k1", "k2"], [SortSpec("ss", True)], 0, 0),              # all-NULL SUM last descending, full order
    (["k1", "k2"], [SortSpec("c", True)], 12, 0),              # COUNT ties -> packed key (k2 nullable, k1 signed)
    (["k1"], [SortSpec("sd", True)], 4, 0),                    # NaN sums first descending, ±0 sums tie
    (["k1", "k2"], [SortSpec("mxe"), SortSpec("k2", True)], 30, 3),  # multi-item, NULL MAX, OFFSET
    (["k2", "k1"], [SortSpec("k2", True), SortSpec("am")], 0, 0),    # nullable key DESC, DECIMAL AVG scaled
    (["k1", "k2"], [SortSpec("ad", True), SortSpec("mne")], 25, 5),  # DOUBLE AVG with NaN, signed-zero MIN
    (["k2"], [SortSpec("si")], 0, 0),                          # INT SUM, NULL key group
]


def order_request(case):
    keys, order, limit, offset = case
    return Request(keys, ORDER_AGGS, order, limit, offset)


def check_order(oracle_lib, t, req, dop=2):
    want = brute_order(t, req)
    got = rows_of(oracle_lib.run(t, req, dop))
    assert len(got) == len(want), (len(got), len(want))
    for i, (a, b) in enumerate(zip(got, want)):
        assert same_values(a, b), f"position {i}: {a} != {b}"


@pytest.mark.parametrize("case", range(len(ORDER_CASES)))
def test_order_against_brute_force(oracle_lib, case):
    for wide in (False, True):
        check_order(oracle_lib, order_table(seed=case, wide=wide), order_request(ORDER_CASES[case]))


# each planted mistake and a case that must expose it
PLANTS = [("DECIMAL_LOW_WORD", 0), ("AVG_TRUNC", 1), ("NULL_AS_ZERO", 2), ("NULL_AS_ZERO", 3),
          ("TIES_BY_GID", 4), ("NAN_FIRST", 5), ("TIES_SIGNED", 4)]


@pytest.mark.parametrize("plant,case", PLANTS)
def test_planted_comparator_mistakes_fail(oracle_lib, plant, case):
    t = order_table(seed=case)
    req = order_request(ORDER_CASES[case])
    check_order(oracle_lib, t, req)  # the true comparator passes
    with oracle_lib.planted(getattr(oracle_lib.planted, plant)):
        with pytest.raises(AssertionError):
            check_order(oracle_lib, t, req)


# ---- 128-bit keys ------------------------------------------------------------------------

WIDE_AGGS = [AggSpec("count", None, "c"), AggSpec("sum", "v", "sv"), AggSpec("min", "w", "mnw"),
             AggSpec("max", "w", "mxw"), AggSpec("avg", "d", "ad"), AggSpec("sum", "x", "sx"),
             AggSpec("sum", "w", "sw")]

WIDE_CASES = [
    (["a", "b"], [SortSpec("c", True)], 60, 0),       # COUNT ties over keys of both signs and the extremes
    (["a", "b"], [SortSpec("b", True), SortSpec("a")], 40, 5),  # key items over the full range
    (["b", "a"], [SortSpec("sx")], 50, 0),            # NULL SUMs first, ties by the packed key
    (["a", "c"], [SortSpec("sw", True)], 30, 0),      # wrapped BIGINT SUM, nullable INT key
    (["c", "a"], [SortSpec("c"), SortSpec("mxw", True)], 45, 2),
]


@pytest.mark.parametrize("case", range(len(WIDE_CASES)))
def test_wide_keys_against_brute_force(oracle_lib, case):
    """Keys that pack into up to 128 bits (two full-range BIGINTs; a BIGINT
    and a nullable INT) under the brute-force comparator."""
    keys, order, limit, offset = WIDE_CASES[case]
    t = wide_key_table(n=5000, groups=1200, seed=case, nullable="c" in keys)
    check_order(oracle_lib, t, Request(keys, WIDE_AGGS, order, limit, offset))


def test_planted_tie_mistake_fails_on_wide_keys(oracle_lib):
    t = wide_key_table(n=5000, groups=1200, seed=0)
    keys, order, limit, _ = WIDE_CASES[0]
    req = Request(keys, WIDE_AGGS, order, limit)
    check_order(oracle_lib, t, req)
    with oracle_lib.planted(oracle_lib.planted.TIES_SIGNED):
        with pytest.raises(AssertionError):
            check_order(oracle_lib, t, req)


def test_min_max_double_nan_and_signed_zero(oracle_lib):
    """MIN/MAX of DOUBLE over NaNs and signed zeros: the reference's result
    depends on row order (the first value seeds, IEEE compares), so the
    oracle's must be one it can produce for some order."""
    import struct
    from helpers import nan_zero_table, reachable_extremes
    t = nan_zero_table()
    req = Request(["k"], [AggSpec("min", "x", "mn"), AggSpec("max", "x", "mx")], [SortSpec("k")])
    for dop in (1, 4):
        for k, mn, mx in rows_of(oracle_lib.run(t, req, dop)):
            xs = [float(v) for v in t.columns["x"][t.columns["k"] == k]]
            bits = lambda v: struct.unpack("<Q", struct.pack("<d", v))[0]
            assert bits(mn) in reachable_extremes(xs, False), (k, mn)
            assert bits(mx) in reachable_extremes(xs, True), (k, mx)
