"""HAVING on the CPU: the oracle's group filter pinned against brute force,
the reference's evaluation rules as invariants, and the operator's
translation of a HAVING tree into the C-ABI's postfix nodes.

The rules (src/exec/v2_shape_group_topn.zig:2385-2535): a leaf compares an
output column with a literal, integers exactly and anything involving a
double in f64 with NaN largest; a NULL output fails the leaf; NOT is
two-valued (NOT of a failed-by-NULL leaf passes); groups are filtered before
ORDER BY / OFFSET / LIMIT.
"""
import math
from decimal import Decimal
from fractions import Fraction

import numpy as np
import pytest

from helpers import ALL_AGGS, brute_force, random_table, rows_of
from thindb_amd import gpu
from thindb_amd.exec import (And, AggSpec, Column, Declined, IsNotNull, IsNull, Not, Or, Predicate, Request,
                             SortSpec, Table, _having_postfix, output_schema)
from thindb_amd.gpu import HavingKind, PredOp, Type

OPS = {"eq": lambda o: o == 0, "neq": lambda o: o != 0, "lt": lambda o: o < 0, "lte": lambda o: o <= 0,
       "gt": lambda o: o > 0, "gte": lambda o: o >= 0}


def _num(x):
    """An exact number (Fraction) or a float; NaN stays a float."""
    if isinstance(x, float):
        return x if math.isnan(x) else Fraction(x)
    if isinstance(x, Decimal):
        return Fraction(x)
    return Fraction(int(x))


def _order(a, b) -> int:
    if isinstance(a, float) and math.isnan(a):
        return 0 if isinstance(b, float) and math.isnan(b) else 1
    if isinstance(b, float) and math.isnan(b):
        return -1
    return (a > b) - (a < b)


def passes(expr, row: dict, scales: dict) -> bool:
    """The HAVING tree over one output row ({name: value or None}); DECIMAL
    values are mantissas, so a literal compares at the column's scale."""
    if isinstance(expr, Predicate):
        v = row[expr.col.lower()]
        if v is None:
            return False
        lit = expr.val
        if scales.get(expr.col.lower()):
            lit = Decimal(repr(lit) if isinstance(lit, float) else lit).scaleb(scales[expr.col.lower()])
        return OPS[expr.op](_order(_num(v), _num(lit)))
    if isinstance(expr, IsNull):
        return row[expr.col.lower()] is None
    if isinstance(expr, IsNotNull):
        return row[expr.col.lower()] is not None
    if isinstance(expr, Not):
        return not passes(expr.child, row, scales)
    if isinstance(expr, And):
        return all(passes(c, row, scales) for c in expr.children)
    return any(passes(c, row, scales) for c in expr.children)


def _table():
    """random_table with the groups k1 in [-50, -46) holding no non-null `s`
    and `b` (so MIN(s), SUM(b) are NULL there)."""
    rng = np.random.default_rng(21)
    t = random_table(rng, 40_000, rg_rows=997, nulls=True)
    dead = t.columns["k1"] < -46
    t.validity["s"] = t.validity["s"] & ~dead
    t.validity["b"] = t.validity["b"] & ~dead
    t._packed.clear()
    return t


def _names(t, req):
    return [c.name.lower() for c in output_schema(t, req)]


def _scales(t, req):
    return {c.name.lower(): c.scale for c in output_schema(t, req) if c.type in (Type.DECIMAL64, Type.DECIMAL128)}


def _mid(values, q):
    """A literal strictly between two neighbouring distinct values near the
    q-quantile, so float summation order cannot move a group across it."""
    v = sorted({x for x in values if x is not None})
    i = max(1, min(len(v) - 1, int(q * len(v))))
    return (v[i - 1] + v[i]) / 2


def _cases(t):
    base = Request(["k1"], ALL_AGGS)
    bf = brute_force(t, base)
    col = lambda name: [vals[[a.as_ for a in ALL_AGGS].index(name)] for vals in bf.values()]  # noqa: E731
    c_mid = int(_mid(col("c"), 0.5))
    return [
        Predicate("c", "gt", c_mid),
        And((Predicate("sb", "lt", 0), Predicate("ai", "gte", _mid(col("ai"), 0.3)))),
        Or((IsNull("mins"), Predicate("sd", "gt", _mid(col("sd"), 0.8)))),
        Not(And((Predicate("c", "gte", c_mid), Predicate("k1", "lt", 10)))),
        Predicate("sm", "gte", Decimal(_mid(col("sm"), 0.4)) / 100),
        And((Predicate("k1", "neq", 3), Predicate("minf", "lt", float(_mid(col("minf"), 0.6))),
             Predicate("am", "lt", float(_mid(col("am"), 0.7))))),
        Predicate("cs", "lte", c_mid / 2 + 0.5),
        IsNotNull("sb"),
        Not(Predicate("sb", "gt", 0)),       # NULL SUM fails the leaf, NOT passes it
        Predicate("maxb", "eq", sorted(x for x in col("maxb") if x is not None)[3]),
        Or((Predicate("mint", "gt", 10 ** 12), Predicate("ad", "neq", _mid(col("ad"), 0.5)))),
        And(()),                              # arity 0: every group
        Or(()),                               # arity 0: none
    ]


def test_oracle_having_matches_brute_force(oracle_lib):
    """For each HAVING tree, the oracle keeps exactly the groups whose
    brute-force aggregates (exact integers, fsum) pass the tree."""
    t = _table()
    bf = brute_force(t, Request(["k1"], ALL_AGGS))
    for having in _cases(t):
        req = Request(["k1"], ALL_AGGS, having_filter=having)
        names = _names(t, req)
        want = sorted(k for k, vals in bf.items() if passes(having, dict(zip(names, list(k) + vals)), _scales(t, req)))
        got = sorted(r[:1] for r in rows_of(oracle_lib.run(t, req, dop=3)))
        assert got == want, having
    # the dead groups exist and are NULL where expected
    dead = [k for k, vals in bf.items() if k[0] < -46]
    assert dead and all(bf[k][[a.as_ for a in ALL_AGGS].index("mins")] is None for k in dead)


def test_oracle_having_before_order_and_limit(oracle_lib):
    """HAVING then ORDER BY / OFFSET / LIMIT equals the ordered result of
    every group, filtered, then sliced."""
    t = _table()
    order = [SortSpec("c", True), SortSpec("k1")]
    for having in _cases(t)[:6]:
        every = oracle_lib.run(t, Request(["k1", "k2"], ALL_AGGS, order), dop=1)
        names = [c.name.lower() for c in output_schema(t, Request(["k1", "k2"], ALL_AGGS))]
        sc = _scales(t, Request(["k1", "k2"], ALL_AGGS))
        kept = [r for r in rows_of(every) if passes(having, dict(zip(names, r)), sc)]
        for limit, offset in ((0, 0), (7, 0), (5, 3), (10 ** 6, 2)):
            req = Request(["k1", "k2"], ALL_AGGS, order, limit=limit, offset=offset, having_filter=having)
            got = rows_of(oracle_lib.run(t, req, dop=1))
            end = len(kept) if limit == 0 else offset + limit
            assert got == kept[offset:end], (having, limit, offset)


def test_oracle_having_invariants(oracle_lib):
    """p and NOT p split the groups; p AND NOT p keeps none; De Morgan."""
    t = _table()
    n_all = oracle_lib.run(t, Request(["k1", "k2"], ALL_AGGS), dop=2).n_rows
    count = lambda h: oracle_lib.run(t, Request(["k1", "k2"], ALL_AGGS, having_filter=h), dop=2).n_rows  # noqa
    for p in _cases(t)[:-2]:
        q = Predicate("c", "lt", 60)
        assert count(p) + count(Not(p)) == n_all
        assert count(And((p, Not(p)))) == 0
        assert count(Or((p, Not(p)))) == n_all
        assert count(Not(And((p, q)))) == count(Or((Not(p), Not(q))))
        assert count(And((p, q))) + count(And((p, Not(q)))) == count(p)


def test_oracle_having_wrapped_sum_and_wide_literals(oracle_lib):
    """A BIGINT SUM compares by its wrapped 64-bit value; a DECIMAL128 SUM
    against literals beyond 64 bits; COUNT against a double literal."""
    n = 3000
    k = np.repeat(np.arange(3, dtype=np.int32), n // 3)
    v = np.full(n, 2 ** 62, np.int64)
    v[k == 1] = -5
    m = np.full(n, 9 * 10 ** 17, np.int64)
    m[k == 1] = -7
    m[k == 2] = 1
    t = Table([Column("k", Type.INT), Column("v", Type.BIGINT), Column("m", Type.DECIMAL64, False, 3)],
              {"k": k, "v": v, "m": m})
    aggs = [AggSpec("sum", "v", "s"), AggSpec("sum", "m", "sm"), AggSpec("count", None, "c")]
    wrapped = ((1000 * 2 ** 62 + 2 ** 63) % 2 ** 64) - 2 ** 63
    assert wrapped == 0  # 1000 * 2^62 is a multiple of 2^64
    got = lambda h: sorted(r[0] for r in rows_of(oracle_lib.run(t, Request(["k"], aggs, having_filter=h))))  # noqa
    assert got(Predicate("s", "eq", 0)) == [0, 2]
    assert got(Predicate("s", "lt", 0)) == [1]
    # 1000 * 9e17 mantissa = 9e20 > 2^64: compared at full width
    assert got(Predicate("sm", "gt", Decimal(2 ** 64) / 1000)) == [0]
    assert got(Predicate("sm", "eq", Decimal(9 * 10 ** 17))) == [0]     # mantissa 9e20 at scale 3
    assert got(Predicate("c", "gt", 999.5)) == [0, 1, 2]
    assert got(Predicate("c", "lt", 1000.0)) == []


def test_having_postfix_translation():
    """The operator's HAVING tree -> postfix nodes, literals in the output
    column's domain, and the forms it declines."""
    t = Table([Column("k", Type.INT), Column("m", Type.DECIMAL64, True, 2), Column("d", Type.DOUBLE)],
              {"k": np.zeros(4, np.int32), "m": np.zeros(4, np.int64), "d": np.zeros(4)})
    req = Request(["k"], [AggSpec("sum", "m", "sm"), AggSpec("avg", "d", "ad"), AggSpec("count", None, "c")])
    schema = output_schema(t, req)
    out = []
    _having_postfix(Or((Not(Predicate("sm", "gte", Decimal("1.5"))), And((Predicate("c", "gt", 2.5),
                                                                             IsNull("sm"),
                                                                             Predicate("ad", "lt", 3))))),
                    schema, out)
    assert out == [(HavingKind.CMP, 1, PredOp.GTE, 150), (HavingKind.NOT,),
                   (HavingKind.CMP, 3, PredOp.GT, 2.5), (HavingKind.IS_NULL, 1),
                   (HavingKind.CMP, 2, PredOp.LT, 3), (HavingKind.AND, 3), (HavingKind.OR, 2)]
    # a DECIMAL literal below the scale compares as a double mantissa
    out = []
    _having_postfix(Predicate("sm", "lt", Decimal("0.125")), schema, out)
    assert out == [(HavingKind.CMP, 1, PredOp.LT, 12.5)]
    for bad in (Predicate("nope", "gt", 1), Predicate("c", "like", 1), "text"):
        with pytest.raises(Declined):
            _having_postfix(bad, schema, [])
    # the C-ABI node: an i128 literal as two halves
    h = gpu.having_node((HavingKind.CMP, 1, PredOp.EQ, -(2 ** 100)))
    assert (h.lit_hi, h.lit_lo, h.lit_float) == (-(2 ** 36), 0, 0)
    h = gpu.having_node((HavingKind.CMP, 1, PredOp.EQ, 2 ** 64 + 7))
    assert (h.lit_hi, h.lit_lo) == (1, 7)
    with pytest.raises(gpu.Unsupported):
        gpu.having_node((HavingKind.CMP, 1, PredOp.EQ, 2 ** 127))
