"""WHERE trees (AND / OR / NOT / IS [NOT] NULL over `col op literal`
leaves) on the CPU: the oracle against brute force.

Semantics (evaluatePredicate, src/exec/predicate.zig:1727-1790): a
comparison with a NULL value does not match, IS [NOT] NULL read the
validity, NOT inverts its child's matches (the SQL front end pushes NOT
into the leaves first, src/sql/parse_predicate.zig:143-170, so a NOT that
reaches the operator is this two-valued inversion), AND / OR combine.
"""
from decimal import Decimal

import numpy as np

from helpers import ALL_AGGS, brute_force, random_table, rows_of
from thindb_amd.exec import And, IsNotNull, IsNull, Not, Or, Predicate, Request, Table

OPS = {"eq": lambda a, b: a == b, "neq": lambda a, b: a != b, "lt": lambda a, b: a < b,
       "lte": lambda a, b: a <= b, "gt": lambda a, b: a > b, "gte": lambda a, b: a >= b}


def row_mask(t: Table, expr) -> np.ndarray:
    """Per-row matches of `expr`, written from the definition."""
    n = t.n_rows
    if isinstance(expr, Predicate):
        c = t.column(expr.col)
        vals = t.columns[c.name]
        lit = expr.val
        if c.scale:
            lit = Decimal(repr(lit) if isinstance(lit, float) else lit).scaleb(c.scale)
        valid = t.validity.get(c.name, np.ones(n, bool))
        x = [(float(v) if vals.dtype.kind == "f" else int(v)) for v in vals]
        return np.array([bool(valid[r]) and OPS[expr.op](x[r], lit) for r in range(n)])
    if isinstance(expr, IsNull):
        return ~t.validity.get(t.column(expr.col).name, np.ones(n, bool))
    if isinstance(expr, IsNotNull):
        return t.validity.get(t.column(expr.col).name, np.ones(n, bool)).copy()
    if isinstance(expr, Not):
        return ~row_mask(t, expr.child)
    if isinstance(expr, And):
        m = np.ones(n, bool)
        for ch in expr.children:
            m &= row_mask(t, ch)
        return m
    m = np.zeros(n, bool)
    for ch in expr.children:
        m |= row_mask(t, ch)
    return m


def subset(t: Table, mask: np.ndarray) -> Table:
    return Table(t.schema, {c.name: t.columns[c.name][mask] for c in t.schema},
                 {k: np.asarray(v)[mask] for k, v in t.validity.items()}, rg_rows=t.rg_rows)


TREES = [
    Or((Predicate("k2", "eq", 1), Predicate("i", "lt", 0))),
    Not(Predicate("s", "gt", 100)),
    And((Predicate("i", "gt", -2 ** 30), Or((IsNull("d"), Predicate("d", "gte", 2500.0))))),
    Or((And((Predicate("k1", "lt", 0), Predicate("f", "gt", 0.5))),
        Not(Or((Predicate("m", "lte", 0), IsNotNull("b")))))),
    And((Not(IsNull("s")), Or((Predicate("m", "gt", 5e9), Predicate("t", "lt", 10 ** 11), Predicate("k2", "neq", 0))))),
    Or(()),
]


def test_oracle_where_trees_match_brute_force(oracle_lib):
    rng = np.random.default_rng(31)
    t = random_table(rng, 12_000, rg_rows=701, nulls=True)
    for tree in TREES:
        mask = row_mask(t, tree)
        want = brute_force(subset(t, mask), Request(["k1", "k2"], ALL_AGGS)) if mask.any() else {}
        got = rows_of(oracle_lib.run(t, Request(["k1", "k2"], ALL_AGGS, where_filter=tree), dop=2))
        assert len(got) == len(want), tree
        for r in got:
            w = want[tuple(r[:2])]
            for j, (x, y) in enumerate(zip(r[2:], w)):
                if isinstance(y, float) and y is not None and x is not None:
                    assert abs(x - y) <= 1e-12 * max(1.0, abs(y)), (tree, r[:2], j)
                else:
                    assert x == y, (tree, r[:2], j, x, y)
    # NOT is the complement of its child on every row: |p| + |NOT p| = rows
    for tree in TREES[:5]:
        a = sum(r[1] for r in rows_of(oracle_lib.run(t, Request(["k2"], ALL_AGGS[:1], where_filter=tree), 1)))
        b = sum(r[1] for r in rows_of(oracle_lib.run(t, Request(["k2"], ALL_AGGS[:1], where_filter=Not(tree)), 1)))
        assert a + b == t.n_rows
