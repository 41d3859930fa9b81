"""CPU tests of the host-side operator logic (request translation, literal
coercion, schema, table cutting) and of the generator's distributions."""
import math
from decimal import Decimal

import numpy as np
import pytest

from thindb_amd import exec as ex
from thindb_amd.exec import And, AggSpec, Column, Not, Or, Predicate, Request, Table
from thindb_amd.gpu import PredOp, Type


def test_literal_coercion_matches_integer_semantics():
    c = Column("x", Type.INT)
    assert ex._coerce(c, "lt", 2.5) == (PredOp.LT, 3)     # x < 2.5  <=> x < 3
    assert ex._coerce(c, "lte", 2.5) == (PredOp.LTE, 2)   # x <= 2.5 <=> x <= 2
    assert ex._coerce(c, "gt", 2.5) == (PredOp.GT, 2)     # x > 2.5  <=> x > 2
    assert ex._coerce(c, "gte", 2.5) == (PredOp.GTE, 3)   # x >= 2.5 <=> x >= 3
    assert ex._coerce(c, "gt", -2.5) == (PredOp.GT, -3)
    assert ex._coerce(c, "eq", 7) == (PredOp.EQ, 7)
    with pytest.raises(ex.Declined):
        ex._coerce(c, "eq", 2.5)
    d = Column("m", Type.DECIMAL64, scale=2)
    assert ex._coerce(d, "gte", Decimal("1.25")) == (PredOp.GTE, 125)
    assert ex._coerce(d, "lt", 3) == (PredOp.LT, 300)
    f = Column("p", Type.DOUBLE)
    assert ex._coerce(f, "lt", 3) == (PredOp.LT, 3.0)
    # exhaustive check of the integer rounding against the definition
    for lit in [x / 4 for x in range(-20, 21)]:
        for op in ("lt", "lte", "gt", "gte"):
            o, b = ex._coerce(c, op, lit)
            for v in range(-8, 9):
                want = {"lt": v < lit, "lte": v <= lit, "gt": v > lit, "gte": v >= lit}[op]
                got = {PredOp.LT: v < b, PredOp.LTE: v <= b, PredOp.GT: v > b, PredOp.GTE: v >= b}[o]
                assert want == got, (op, lit, v)


def test_where_conjuncts_and_tree():
    """Top-level AND leaves become fused predicate terms; everything else one
    postfix tree AND-ed with them."""
    p, q = Predicate("a", "gt", 1), Predicate("b", "lt", 2)
    assert ex._conjuncts(And((p, And((q,))))) == [p, q]
    assert ex._conjuncts(Or((p, q))) == [Or((p, q))]
    t = Table([Column("a", Type.INT), Column("b", Type.DOUBLE, True)], {"a": [1], "b": [1.0]}, {"b": [True]})
    out, cols = [], []
    ex._where_postfix(Or((Not(p), And((q, ex.IsNull("b"))))), t, out, cols)
    K = ex.gpu.HavingKind
    assert out == [(K.CMP, 0, Type.INT, PredOp.GT, 1), (K.NOT,), (K.CMP, 1, Type.DOUBLE, PredOp.LT, 2.0),
                   (K.IS_NULL, 1), (K.AND, 2), (K.OR, 2)]
    assert cols == ["a", "b", "b"]
    with pytest.raises(ex.Declined):
        ex._where_postfix("text", t, [], [])


def test_output_schema_types():
    t = Table([Column("k", Type.INT), Column("v", Type.BIGINT, True), Column("m", Type.DECIMAL64, False, 3),
               Column("d", Type.FLOAT)], {"k": [1], "v": [2], "m": [3], "d": [1.0]}, {"v": [True]})
    req = Request(["k"], [AggSpec("count", None, "c"), AggSpec("count", "v", "cv"), AggSpec("sum", "v", "sv"),
                          AggSpec("sum", "m", "sm"), AggSpec("avg", "m", "am"), AggSpec("min", "m", "mm"),
                          AggSpec("sum", "d", "sd"), AggSpec("max", "d", "xd")])
    s = ex.output_schema(t, req)
    assert [(c.name, c.type, c.nullable, c.scale) for c in s] == [
        ("k", Type.INT, False, 0), ("c", Type.BIGINT, False, 0), ("cv", Type.BIGINT, False, 0),
        ("sv", Type.BIGINT, True, 0), ("sm", Type.DECIMAL128, True, 3), ("am", Type.DOUBLE, True, 0),
        ("mm", Type.DECIMAL64, True, 3), ("sd", Type.DOUBLE, True, 0), ("xd", Type.FLOAT, True, 0)]


def test_table_cuts_segments_and_row_groups():
    n = 10_000
    t = Table([Column("a", Type.INT)], {"a": np.arange(n)}, rg_rows=1000, segment_rows=3500)
    starts = [(rg.segment_id, rg.idx, rg.start, rg.n_rows) for rg in t.row_groups]
    assert starts[:5] == [(0, 0, 0, 1000), (0, 1, 1000, 1000), (0, 2, 2000, 1000), (0, 3, 3000, 500),
                          (1, 0, 3500, 1000)]
    assert sum(s[3] for s in starts) == n
    vals, bm = t.block("a", t.row_groups[4])
    assert vals[0] == 3500 and bm is None


def test_validity_packing_is_lsb_first():
    t = Table([Column("a", Type.INT, True)], {"a": np.arange(10)},
              {"a": np.array([1, 0, 1, 1, 0, 0, 0, 1, 0, 1], bool)}, rg_rows=16)
    _, bm = t.block("a", t.row_groups[0])
    # bit r at byte r>>3, bit r&7 (src/storage/column.zig:207-210)
    assert list(bm) == [0b10001101, 0b10]


def test_generator_distributions(gen_lib):
    g = gen_lib
    n = 2_000_000
    f = g.filtered(n, seed=2)
    sel = ((f["ts"] >= g.FILTER_LO) & (f["ts"] < g.FILTER_HI)).mean()
    assert abs(sel - 7 / 30) < 0.002
    assert f["region"].min() >= 0 and f["region"].max() < 10_000
    # the hottest id's share is Zipf(1.1)'s rank-1 mass
    top = np.bincount(f["region"], minlength=10_000).max() / n
    assert abs(top - g.zipf_pmf(10_000, 1.1, 1)) < 0.003
    assert f["amount"].min() >= 0 and f["amount"].max() <= 10 ** 12
    assert abs(np.median(np.log(np.maximum(f["amount"], 1))) - 8.0) < 0.05
    assert 0 <= f["price"].min() and f["price"].max() < 1000
    # same seed, any offset: identical rows
    part = g.filtered(1000, seed=2, row0=12345)
    for k in f:
        assert np.array_equal(part[k], f[k][12345:13345])
    h = g.highcard(n, seed=3)
    assert h["user_id"].max() < 2 * 10 ** 7 and h["bucket"].max() == 15 and h["v"].max() < 10 ** 5
    assert abs(h["w"].mean()) < 0.01 and abs(h["w"].std() - 1) < 0.01
    p = g.plain(n, seed=1)
    assert p["k"].min() == 0 and p["k"].max() == 1023 and abs(p["v"]).max() <= 10 ** 6
    rows, card, filt = g.batched_shape(7, 3)
    assert 10 ** 4 <= rows <= 2 * 10 ** 7 and 4 <= card <= 10 ** 6


def test_declines_where_over_computed_columns():
    """A request whose WHERE has conjuncts over computed columns
    (residual_filter / computed_where, v2_shape_group_topn.zig:71-76) is
    declined before any device work, so its filter is never dropped."""
    import pytest
    from thindb_amd.exec import AggSpec, Column, Declined, Predicate, Request, Table, build_plan
    from thindb_amd.gpu import Type
    t = Table([Column("k", Type.INT), Column("v", Type.BIGINT)],
              {"k": np.arange(10, dtype=np.int32), "v": np.arange(10, dtype=np.int64)})
    for extra in ({"residual_filter": Predicate("v2", "gt", 0)}, {"computed_where": object()}):
        with pytest.raises(Declined, match="computed columns"):
            build_plan(None, t, Request(["k"], [AggSpec("sum", "v", "s")], **extra))
