"""Shared test helpers: golden tables, a pure-Python brute force, and result
comparison under the reference's tie rule."""
from __future__ import annotations

import dataclasses
import json
import math
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np

from thindb_amd.exec import And, AggSpec, Column, Predicate, Request, SortSpec, Table
from thindb_amd.gpu import Type

GOLDEN = Path(__file__).resolve().parent / "golden"

TYPE_NAMES = {t.name: t for t in Type}
INT_RANGE = {Type.BOOLEAN: (0, 1), Type.TINYINT: (-128, 127), Type.SMALLINT: (-32768, 32767),
             Type.INT: (-2 ** 31, 2 ** 31 - 1), Type.DATE: (-2 ** 31, 2 ** 31 - 1),
             Type.BIGINT: (-2 ** 63, 2 ** 63 - 1), Type.DATETIME: (-2 ** 63, 2 ** 63 - 1),
             Type.DECIMAL64: (-2 ** 63, 2 ** 63 - 1)}


def golden(name: str) -> dict:
    return json.loads((GOLDEN / f"{name}.json").read_text())


def _schema(spec) -> List[Column]:
    out = []
    for s in spec:
        out.append(Column(s[0], TYPE_NAMES[s[1]], bool(s[2]), s[3] if len(s) > 3 else 0))
    return out


def wrap_table() -> Table:
    g = golden("wrap_sum")
    schema = _schema(g["schema"])
    f = g["filler"]
    n = len(g["head"]) + f["rows"]
    ids = np.empty(n, np.int64)
    gs = np.empty(n, np.int32)
    hk = np.empty(n, np.int32)
    v = np.zeros(n, np.int64)
    valid = np.ones(n, bool)
    for i, (a, b, c, d) in enumerate(g["head"]):
        ids[i], gs[i], hk[i] = a, b, c
        if d is None:
            valid[i] = False
        else:
            v[i] = d
    h = len(g["head"])
    ids[h:] = f["id0"] + np.arange(f["rows"])
    gs[h:] = f["g"]
    hk[h:] = f["hk0"] + np.arange(f["rows"])
    v[h:] = f["v"]
    return Table(schema, {"id": ids, "g": gs, "hk": hk, "v": v}, {"v": valid}, rg_rows=g["row_group_size"])


def decimal_table() -> Table:
    g = golden("decimal_group")
    rows = g["rows"]
    return Table(_schema(g["schema"]), {"id": np.array([r[0] for r in rows], np.int32),
                                        "grp": np.array([r[1] for r in rows], np.int32),
                                        "v": np.array([r[2] for r in rows], np.int64)},
                 {"v": np.ones(len(rows), bool)})


def float_table() -> Table:
    g = golden("float_group")
    rows = g["rows"]
    return Table(_schema(g["schema"]), {"id": np.array([r[0] for r in rows], np.int64),
                                        "k": np.array([r[1] for r in rows], np.int16),
                                        "sd": np.array([r[2] for r in rows], np.float64),
                                        "sf": np.array([r[3] for r in rows], np.float32)})


def heavy_table() -> Table:
    g = golden("heavy_groups")
    k = []
    for grp in range(g["heavy_groups"]):
        k += [grp] * (6000 - 250 * grp)
    k += [1000 + j for j in range(g["unique_rows"])]
    k = np.array(k, np.int64)
    total = len(k)
    assert total == g["total"]
    i = np.arange(total)
    v = (i % 7).astype(np.int32)
    ids = (i * 7919) % total
    order = np.argsort(ids, kind="stable")  # physical order is the order key
    schema = [Column("id", Type.BIGINT), Column("k", Type.BIGINT), Column("v", Type.INT)]
    return Table(schema, {"id": ids[order].astype(np.int64), "k": k[order], "v": v[order]},
                 rg_rows=g["row_group_size"])


def ranked_table() -> Table:
    """The ranked-emit test's table (ranked_pages.json): 140,009 groups."""
    g = golden("ranked_pages")
    n_groups = g["group_count"]
    k = np.arange(n_groups, dtype=np.int64)
    reps = 1 + k % 3
    kk = np.repeat(k, reps)
    j = np.arange(len(kk)) - np.repeat(np.cumsum(reps) - reps, reps)
    schema = [Column("id", Type.BIGINT), Column("g", Type.BIGINT), Column("v", Type.INT)]
    return Table(schema, {"id": np.arange(len(kk), dtype=np.int64), "g": kk * 1_000_003,
                          "v": ((kk * 31 + j * 7) % 17).astype(np.int32)}, rg_rows=g["row_group_size"])


def nullable_agg_table() -> Table:
    """The nullable-aggregate test's table (nullable_aggs.json)."""
    g = golden("nullable_aggs")
    G = g["group_count"]
    i = np.arange(6 * G, dtype=np.int64)
    grp = (i // 2) % G
    z = grp == 0
    schema = [Column("id", Type.BIGINT), Column("g", Type.INT, True), Column("v", Type.INT, True),
              Column("b", Type.BIGINT, True), Column("d", Type.DOUBLE, True)]
    cols = {"id": i, "g": grp.astype(np.int32), "v": np.where(z | (i % 7 == 0), 0, (i * 37) % 1000 - 500).astype(np.int32),
            "b": np.where(z | (i % 5 == 0), 0, i * 3_000_000_019), "d": (((i * 13) % 512) - 256.0) * 0.25}
    valid = {"g": np.ones(len(i), bool), "v": ~(z | (i % 7 == 0)), "b": ~(z | (i % 5 == 0)), "d": ~(z | (i % 3 == 0))}
    return Table(schema, cols, valid, rg_rows=g["row_group_size"])


def order_table(seed: int = 0, n: int = 4000, wide: bool = False) -> Table:
    """A table built to separate ORDER BY comparators: keys spanning both
    signs (the packed key's bit order differs from the numeric order), a
    nullable key, DECIMAL sums beyond 64 bits, groups whose SMALLINT input is
    all NULL, integer AVGs sharing integer parts, DOUBLE sums holding NaN and
    signed zeros (values are multiples of 1/4, so every summation order gives
    the same sum), and COUNT(*) ties.  `wide` spreads k1 over a 6e8-value
    span (a sparse key domain: the engine's HBM hash table)."""
    rng = np.random.default_rng(seed)
    k1 = rng.integers(-6, 7, n).astype(np.int32)
    k1_key = k1 * 50_000_001 if wide else k1
    k2 = rng.integers(-2, 3, n).astype(np.int8)
    d = rng.integers(-256, 257, n) * 0.25
    d = np.where(k1 == -5, np.nan, d)
    d = np.where(k1 == 0, np.where(rng.random(n) < 0.5, -0.0, 0.0), d)
    e = rng.integers(-8, 9, n) * 0.5
    e = np.where(e == 0, np.where(rng.random(n) < 0.5, -0.0, 0.0), e)
    schema = [Column("k1", Type.INT), Column("k2", Type.TINYINT, True), Column("m", Type.DECIMAL64, False, 2),
              Column("s", Type.SMALLINT, True), Column("i", Type.INT), Column("d", Type.DOUBLE),
              Column("e", Type.DOUBLE, True)]
    cols = {"k1": k1_key.astype(np.int32), "k2": k2, "m": rng.integers(-9 * 10 ** 18, 9 * 10 ** 18, n, dtype=np.int64),
            "s": rng.integers(-300, 300, n).astype(np.int16), "i": rng.integers(-5, 6, n).astype(np.int32),
            "d": d, "e": e}
    valid = {"k2": rng.random(n) > 0.15, "s": (k1 != 3) & (rng.random(n) > 0.3), "e": (k1 != 2) | (k2 != 1)}
    return Table(schema, cols, valid, rg_rows=256)


def wide_key_table(n: int = 20_000, groups: int = 5_000, seed: int = 0, nullable: bool = False,
                   rg_rows: int = 4096) -> Table:
    """GROUP BY keys no relative packing narrows: `a` and `b` are BIGINTs
    drawn from the whole 64-bit range (both signs, the extremes included),
    so (a, b) packs into 128 bits; `c` is an INT of both signs, NULL in about
    one row in nine when `nullable` (a, b, c then pack into 64 + 64 + 33
    bits, beyond 128: with a, c alone into 97).  Rows pick among `groups`
    (a, b, c) triples with a skew, so groups hold from one row to many;
    aggregate inputs: v INT, w BIGINT (full range), d DOUBLE, x SMALLINT
    (nullable)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.iinfo(np.int64).min, np.iinfo(np.int64).max
    ga = rng.integers(lo, hi, groups, dtype=np.int64, endpoint=True)
    gb = rng.integers(lo, hi, groups, dtype=np.int64, endpoint=True)
    ga[:4] = [lo, hi, -1, 0]
    gb[:4] = [hi, lo, 0, -1]
    gb[4:8] = gb[0]                  # pairs sharing b: ties broken by a
    ga[8:12] = ga[1]                 # pairs sharing a: ties broken by b
    gc = rng.integers(-2 ** 31, 2 ** 31, groups, dtype=np.int64).astype(np.int32)
    pick = np.minimum((rng.pareto(1.2, n) * groups / 50).astype(np.int64), groups - 1)
    pick[:groups] = np.arange(groups)  # every group holds a row
    rng.shuffle(pick)
    schema = [Column("a", Type.BIGINT), Column("b", Type.BIGINT), Column("c", Type.INT, nullable),
              Column("v", Type.INT), Column("w", Type.BIGINT), Column("d", Type.DOUBLE),
              Column("x", Type.SMALLINT, True)]
    cols = {"a": ga[pick], "b": gb[pick], "c": gc[pick], "v": rng.integers(-10 ** 6, 10 ** 6, n).astype(np.int32),
            "w": rng.integers(lo, hi, n, dtype=np.int64, endpoint=True), "d": rng.integers(-4000, 4000, n) * 0.25,
            "x": rng.integers(-500, 500, n).astype(np.int16)}
    valid = {"x": rng.random(n) > 0.2}
    if nullable:
        valid["c"] = rng.random(groups)[pick] > 0.11
    return Table(schema, cols, valid, rg_rows=rg_rows)


def nan_zero_table(n: int = 6000, seed: int = 0) -> Table:
    """A DOUBLE column whose groups hold NaNs (both signs, several payloads),
    +0.0 and -0.0: group 0 only NaNs, 1 NaNs and numbers, 2 only signed
    zeros, 3 zeros and positives, 4 zeros and negatives, 5 plain numbers."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 6, n).astype(np.int32)
    nans = np.array([0x7ff8000000000000, 0xfff8000000000000, 0x7ff0000000000001, 0xfff4000000000abc],
                    dtype=np.uint64).view(np.float64)
    pick = rng.integers(0, 4, n)
    zero = np.where(rng.random(n) < 0.5, -0.0, 0.0)
    num = rng.integers(-1000, 1001, n) * 0.5
    x = np.select([k == 0, (k == 1) & (rng.random(n) < 0.3), k == 2, k == 3, k == 4],
                  [nans[pick], nans[pick], zero, np.where(rng.random(n) < 0.3, zero, np.abs(num) + 1),
                   np.where(rng.random(n) < 0.3, zero, -np.abs(num) - 1)], num)
    return Table([Column("k", Type.INT), Column("x", Type.DOUBLE)], {"k": k, "x": x}, rg_rows=512)


def reachable_extremes(xs: Sequence[float], want_max: bool) -> set:
    """Bit patterns a MIN (or MAX) of DOUBLE values can take under the
    reference's fold for SOME row order: the first value seeds the state and
    a later one replaces it when IEEE `<` (`>`) holds
    (foldExtremeFloat, src/exec/v2_lowcard_group.zig:1318-1347; mergeScalars
    :1981-1992).  A NaN first stays (nothing compares below it); otherwise
    NaNs never enter, and among zeros the first one seen stays."""
    import struct
    bits = lambda v: struct.unpack("<Q", struct.pack("<d", v))[0]
    out = {bits(v) for v in xs if v != v}
    nums = [v for v in xs if v == v]
    if nums:
        m = max(nums) if want_max else min(nums)
        out |= {bits(v) for v in nums if v == m}
    return out


# ---- brute force (pure Python, exact) -------------------------------------------------

def _row_passes(table: Table, leaves: Sequence[Predicate], r: int) -> bool:
    for p in leaves:
        c = table.column(p.col)
        if c.name in table.validity and not table.validity[c.name][r]:
            return False
        x = table.columns[c.name][r]
        lit = p.val
        if c.type == Type.DECIMAL64:
            from decimal import Decimal
            lit = Decimal(repr(lit) if isinstance(lit, float) else lit).scaleb(c.scale)
            x = int(x)
        elif c.type in (Type.FLOAT, Type.DOUBLE):
            x = float(x)
        else:
            x = int(x)
        ok = {"eq": x == lit, "neq": x != lit, "lt": x < lit, "lte": x <= lit, "gt": x > lit, "gte": x >= lit}[p.op]
        if not ok:
            return False
    return True


def brute_force(table: Table, request: Request) -> Dict[tuple, list]:
    """Every group's aggregates, straight from the definitions: exact
    integers, math.fsum for floats, NULL (None) for no non-null input."""
    leaves = []
    if request.where_filter is not None:
        stack = [request.where_filter]
        while stack:
            e = stack.pop()
            if isinstance(e, And):
                stack.extend(e.children)
            else:
                leaves.append(e)
    groups: Dict[tuple, list] = {}
    for r in range(table.n_rows):
        if not _row_passes(table, leaves, r):
            continue
        key = tuple(None if (table.column(k).name in table.validity and not table.validity[table.column(k).name][r])
                    else int(table.columns[table.column(k).name][r]) for k in request.group_cols)
        st = groups.setdefault(key, [[0, []] for _ in request.aggs])
        for a, s in zip(request.aggs, st):
            s[0] += 1  # rows
            if a.col is None:
                continue
            c = table.column(a.col)
            if c.name in table.validity and not table.validity[c.name][r]:
                continue
            x = table.columns[c.name][r]
            s[1].append(float(x) if c.type in (Type.FLOAT, Type.DOUBLE) else int(x))
    out = {}
    for key, st in groups.items():
        vals = []
        for a, (rows, xs) in zip(request.aggs, st):
            f = a.func.lower()
            if f == "count":
                vals.append(rows if a.col is None else len(xs))
                continue
            if f == "count_distinct":  # floats: one +0.0 and one NaN (types.canonicalFloat)
                vals.append(len({"nan" if x != x else (0.0 if x == 0 else x) for x in xs}))
                continue
            if not xs:
                vals.append(None)
                continue
            c = table.column(a.col)
            fl = c.type in (Type.FLOAT, Type.DOUBLE)
            if f == "sum":
                if fl:
                    vals.append(math.fsum(xs))
                elif c.type == Type.DECIMAL64:
                    vals.append(sum(xs))
                else:
                    s = sum(xs) & ((1 << 64) - 1)
                    vals.append(s - (1 << 64) if s >> 63 else s)
            elif f == "avg":
                if fl:
                    vals.append(math.fsum(xs) / len(xs))
                else:
                    vals.append(float(sum(xs)) / float(len(xs)) / (10.0 ** c.scale))
            elif f == "min":
                vals.append(min(xs))
            else:
                vals.append(max(xs))
        out[key] = vals
    return out


_KEY_BITS = {Type.BOOLEAN: 8, Type.TINYINT: 8, Type.SMALLINT: 16, Type.INT: 32, Type.DATE: 32}


def _float_cmp(x: float, y: float) -> int:
    """types.floatOrder: -0 = +0, NaN after every number."""
    xn, yn = x != x, y != y
    if xn or yn:
        return 0 if xn and yn else (1 if xn else -1)
    return (x > y) - (x < y)


def brute_order(table: Table, request: Request) -> List[tuple]:
    """The reference's result rows, from the brute force's groups sorted by a
    comparator written from the silo's definition (rowOrderBefore /
    compareOrderKey / compareAggregateValue, src/exec/v2_shape_group_topn.zig:
    1327-1395): each ORDER BY item with NULL first ascending and last
    descending, floats under types.floatOrder, then the packed group key
    ascending (its fields from the last key part to the first, each as
    (validity, two's-complement bits of the type's width)); then OFFSET and
    LIMIT.  No HAVING."""
    import functools
    assert request.having_filter is None
    groups = brute_force(table, request)
    nk = len(request.group_cols)
    names = [k.lower() for k in request.group_cols] + [a.as_.lower() for a in request.aggs]
    key_bits = [_KEY_BITS.get(table.column(k).type, 64) for k in request.group_cols]
    rows = [tuple(k) + tuple(v) for k, v in groups.items()]

    def item(i, desc, a, b):
        x, y = a[i], b[i]
        if x is None or y is None:
            c = 0 if x is None and y is None else (-1 if x is None else 1)
        elif isinstance(x, float) or isinstance(y, float):
            c = _float_cmp(float(x), float(y))
        else:
            c = (x > y) - (x < y)
        return -c if desc else c

    def cmp(a, b):
        for sp in request.order_specs:
            c = item(names.index(sp.col.lower()), sp.desc, a, b)
            if c:
                return c
        for p in range(nk - 1, -1, -1):
            fa = (0, 0) if a[p] is None else (1, a[p] & ((1 << key_bits[p]) - 1))
            fb = (0, 0) if b[p] is None else (1, b[p] & ((1 << key_bits[p]) - 1))
            if fa != fb:
                return -1 if fa < fb else 1
        return 0

    rows.sort(key=functools.cmp_to_key(cmp))
    start = min(request.offset, len(rows))
    end = len(rows) if not request.limit else min(len(rows), start + request.limit)
    return rows[start:end]


def same_values(a: tuple, b: tuple) -> bool:
    """Row equality with NaN equal to NaN (and -0.0 equal to 0.0)."""
    if len(a) != len(b):
        return False
    for x, y in zip(a, b):
        if isinstance(x, float) and isinstance(y, float) and x != x and y != y:
            continue
        if x != y:
            return False
    return True


# ---- result comparison --------------------------------------------------------------------

def columns_of(result) -> List[tuple]:
    """(values, valid-bool) per column, from a GPU Batch or an OracleResult."""
    out = []
    if hasattr(result, "values") and hasattr(result, "row_count"):  # Batch
        for v in result.values:
            n = len(v.values)
            valid = np.ones(n, bool) if v.validity is None else \
                np.unpackbits(v.validity, bitorder="little")[:n].astype(bool)
            out.append((v.values, valid))
    else:
        for c in result.columns:
            out.append((c.values, c.valid))
    return out


def rows_of(result) -> List[tuple]:
    cols = columns_of(result)
    n = len(cols[0][0]) if cols else 0
    rows = []
    for r in range(n):
        row = []
        for vals, valid in cols:
            if not valid[r]:
                row.append(None)
            else:
                x = vals[r]
                row.append(x.item() if hasattr(x, "item") else x)
        rows.append(tuple(row))
    return rows


def _close(a, b, rel: float) -> bool:
    if a is None or b is None:
        return a is b
    if isinstance(a, float) or isinstance(b, float):
        a, b = float(a), float(b)
        if a == b:
            return True
        return abs(a - b) <= rel * max(abs(a), abs(b))
    return a == b


def assert_same_rows(got: List[tuple], want: List[tuple], float_cols: Sequence[int] = (), rel: float = 0.0,
                     n_keys: int = 1, order_cols: Optional[Sequence[int]] = None,
                     full: Optional[List[tuple]] = None, limited: bool = False) -> None:
    """Compares a result with the oracle's.

    `want` is the oracle's result for the same request; `full` (required
    with ORDER BY or LIMIT/OFFSET) the oracle's rows of every group that
    passes HAVING, unordered and unsliced.  Columns in `float_cols` compare
    within relative tolerance `rel` (a DOUBLE SUM/AVG, whose summation order
    is free), everything else bit-exactly.

    * Every returned row equals its group's row of `full`, keys and every
      aggregate, rows tied at a LIMIT cut included (the reference's own
      parity rule: a returned row is a true group row).
    * With ORDER BY (`order_cols` = the output columns ranked on): the ranked
      values agree position by position.  The order is total - ties break
      by the packed group key, as the reference's silo breaks them
      (rowOrderBefore, src/exec/v2_shape_group_topn.zig:1327-1335) - so when
      no ranked column is a float the rows agree position by position,
      keys included.  When one is, two groups whose float values differ by
      less than the tolerance may trade places, so the rows strictly above
      the last ranked value agree as a set instead.
    * Without ORDER BY and without LIMIT/OFFSET the rows agree as a set; with
      `limited` (LIMIT/OFFSET and no ORDER BY: any groups may be returned)
      only the count and the per-row check above apply."""
    assert len(got) == len(want), f"{len(got)} rows vs {len(want)}"

    def eq_row(a, b):
        return all(_close(x, y, rel) if i in float_cols else x == y for i, (x, y) in enumerate(zip(a, b)))

    def key_of(r):  # NULL keys sort first
        return tuple((x is not None, x if x is not None else 0) for x in r[:n_keys])

    if order_cols is not None or limited:
        assert full is not None, "ORDER BY / LIMIT results are checked against the oracle's full group table"
    if full is not None:
        by_key = {tuple(r[:n_keys]): r for r in full}
        assert len(by_key) == len(full), "oracle groups are not unique by key"
        for r in got:
            k = tuple(r[:n_keys])
            assert k in by_key, f"returned group {k} does not exist"
            assert eq_row(r, by_key[k]), f"row {r} != its group {by_key[k]}"
        assert len({tuple(r[:n_keys]) for r in got}) == len(got), "a group is returned twice"
    if order_cols is None:
        if limited:
            return
        ga = sorted(got, key=key_of)
        wa = sorted(want, key=key_of)
        for a, b in zip(ga, wa):
            assert eq_row(a, b), f"{a} != {b}"
        return

    def same_rank(a, b):
        return all(_close(a[i], b[i], rel) if i in float_cols else a[i] == b[i] for i in order_cols)

    for i, (a, b) in enumerate(zip(got, want)):
        assert same_rank(a, b), f"position {i}: ranked value {a} vs {b}"
    if not any(i in float_cols for i in order_cols):
        for i, (a, b) in enumerate(zip(got, want)):
            assert eq_row(a, b), f"position {i}: {a} != {b}"
        return
    if not got:
        return
    last = want[-1]
    above_g = [r for r in got if not same_rank(r, last)]
    above_w = [r for r in want if not same_rank(r, last)]
    assert len(above_g) == len(above_w), (len(above_g), len(above_w))
    by_key_w = {r[:n_keys]: r for r in above_w}
    for r in above_g:
        assert r[:n_keys] in by_key_w and eq_row(r, by_key_w[r[:n_keys]]), f"row {r} not in expected"


def float_cols_of(table: Table, req: Request) -> List[int]:
    """Output columns compared within tolerance: DOUBLE SUM and every AVG."""
    nk = len(req.group_cols)
    out = []
    for j, a in enumerate(req.aggs):
        if a.col is None:
            continue
        c = table.column(a.col)
        if a.func.lower() == "avg" or (a.func.lower() == "sum" and c.type in (Type.FLOAT, Type.DOUBLE)):
            out.append(nk + j)
    return out


def order_cols_of(schema_names: Sequence[str], req: Request) -> Optional[List[int]]:
    if not req.order_specs:
        return None
    names = [n.lower() for n in schema_names]
    return [names.index(s.col.lower()) for s in req.order_specs]


def unsliced(req: Request) -> Request:
    """The same request with no ORDER BY, LIMIT or OFFSET: every group."""
    return dataclasses.replace(req, order_specs=(), limit=0, offset=0)


def check_with_oracle(got: List[tuple], oracle_lib, table: Table, req: Request, dop: int = 2,
                      rel: float = 1e-12, want: Optional[List[tuple]] = None) -> List[tuple]:
    """`got` (rows of a GPU result for `req` over `table`) against the
    oracle under assert_same_rows; returns the oracle's rows."""
    if want is None:
        want = rows_of(oracle_lib.run(table, req, dop))
    names = [k for k in req.group_cols] + [a.as_ for a in req.aggs]
    order_cols = order_cols_of(names, req)
    limited = bool(req.limit or req.offset)
    full = rows_of(oracle_lib.run(table, unsliced(req), dop)) if (order_cols is not None or limited) else None
    assert_same_rows(got, want, float_cols_of(table, req), rel, len(req.group_cols), order_cols, full,
                     limited=limited and order_cols is None)
    return want


def assert_rows_in_groups(got: List[tuple], groups: Dict[tuple, list], n_keys: int, float_cols: Sequence[int] = (),
                          rel: float = 0.0) -> None:
    """Every returned row equals its group's brute-force aggregates."""
    for r in got:
        key = tuple(r[:n_keys])
        assert key in groups, f"group {key} does not exist"
        want = groups[key]
        for j, (x, y) in enumerate(zip(r[n_keys:], want)):
            i = n_keys + j
            if i in float_cols:
                assert _close(x, y, rel), f"group {key} col {i}: {x} vs {y}"
            else:
                assert x == y, f"group {key} col {i}: {x} vs {y}"


def random_table(rng: np.random.Generator, n: int, rg_rows: int = 1000, nulls: bool = True) -> Table:
    """Mixed-type table for the brute-force and parity tests."""
    schema = [Column("k1", Type.INT), Column("k2", Type.TINYINT), Column("b", Type.BIGINT, nulls),
              Column("i", Type.INT), Column("s", Type.SMALLINT, nulls), Column("d", Type.DOUBLE, nulls),
              Column("f", Type.FLOAT), Column("m", Type.DECIMAL64, False, 2), Column("t", Type.DATETIME)]
    cols = {
        "k1": rng.integers(-50, 50, n).astype(np.int32),
        "k2": rng.integers(-3, 4, n).astype(np.int8),
        "b": rng.integers(-2 ** 62, 2 ** 62, n, dtype=np.int64),
        "i": rng.integers(-2 ** 31, 2 ** 31 - 1, n).astype(np.int32),
        "s": rng.integers(-30000, 30000, n).astype(np.int16),
        "d": rng.uniform(-1000.0, 5000.0, n),
        "f": rng.normal(0, 10, n).astype(np.float32),
        "m": rng.integers(-10 ** 12, 10 ** 12, n, dtype=np.int64),
        "t": rng.integers(0, 10 ** 12, n, dtype=np.int64),
    }
    validity = {}
    if nulls:
        validity = {"b": rng.random(n) > 0.1, "s": rng.random(n) > 0.5, "d": rng.random(n) > 0.2}
    return Table(schema, cols, validity, rg_rows=rg_rows)


ALL_AGGS = [AggSpec("count", None, "c"), AggSpec("count", "s", "cs"), AggSpec("sum", "b", "sb"),
            AggSpec("sum", "i", "si"), AggSpec("avg", "i", "ai"), AggSpec("min", "s", "mins"),
            AggSpec("max", "b", "maxb"), AggSpec("sum", "d", "sd"), AggSpec("avg", "d", "ad"),
            AggSpec("min", "f", "minf"), AggSpec("max", "d", "maxd"), AggSpec("sum", "m", "sm"),
            AggSpec("avg", "m", "am"), AggSpec("min", "t", "mint")]


# ---- encoded blocks as thinDB's writer lays them out ----------------------------------------
# src/storage/segment_writer.zig: forEncode (:1410-1470: base = min non-null value, codes
# value - base, NULL or out-of-range placeholders clamped into [0, span]) and rleEncode
# (:1290-1370: runs over the stored stream, NULL placeholders included).  The widths here
# are forced by the caller so that every width the reader accepts is exercised.

def bitmap_bytes(validity) -> bytes:
    """Validity as the reference's bitmap: bit r of byte r / 8 set = row r valid."""
    return np.packbits(np.asarray(validity, dtype=bool), bitorder="little").tobytes()


def for_payload(values: np.ndarray, validity=None, width: int = 0) -> bytes:
    v = np.asarray(values).astype(np.int64)
    ok = np.ones(len(v), bool) if validity is None else np.asarray(validity, bool)
    lo = int(v[ok].min()) if ok.any() else 0
    hi = int(v[ok].max()) if ok.any() else 0
    span = hi - lo
    if not width:
        width = 1 if span < 1 << 8 else 2 if span < 1 << 16 else 4 if span < 1 << 32 else 8
    d = np.where(~ok | (v < lo), 0, np.where(v > hi, span, v.astype(object) - lo)).astype(object)
    codes = b"".join(int(x).to_bytes(width, "little") for x in d)
    head = lo.to_bytes(16, "little", signed=True) + bytes([width, 0, 0, 0])
    return (bitmap_bytes(ok) if validity is not None else b"") + head + codes


def rle_payload(values: np.ndarray, validity=None, lengths=None) -> bytes:
    """`lengths` (optional) replaces the run lengths, to build blocks whose runs
    do not sum to the row count."""
    v = np.asarray(values)
    runs_v, runs_n = [], []
    for x in v.tolist():
        if runs_v and runs_v[-1] == x:
            runs_n[-1] += 1
        else:
            runs_v.append(x)
            runs_n.append(1)
    if lengths is not None:
        runs_n = list(lengths)
    w = v.dtype.itemsize
    body = len(runs_v).to_bytes(4, "little") + bytes([w, 0, 0, 0])
    body += np.asarray(runs_v, dtype=v.dtype).tobytes() + np.asarray(runs_n, dtype="<u4").tobytes()
    return (bitmap_bytes(validity) if validity is not None else b"") + body


def decode_payload(kind: str, payload: bytes, n: int, dtype, has_nulls: bool) -> np.ndarray:
    """Pure-Python decode following decodeForColumn / fillRuns
    (src/storage/segment_reader.zig:1989-2075, :2227-2247): the test's own
    check of the encoders."""
    p = payload[(n + 7) // 8:] if has_nulls else payload
    out = np.zeros(n, dtype=dtype)
    bits = 8 * np.dtype(dtype).itemsize
    if kind == "for":
        base = int.from_bytes(p[0:16], "little", signed=True)
        w = p[16]
        for r in range(n):
            x = (base + int.from_bytes(p[20 + r * w:20 + (r + 1) * w], "little")) % (1 << bits)
            out[r] = x - (1 << bits) if (np.issubdtype(dtype, np.signedinteger) and x >= 1 << (bits - 1)) else x
        return out
    n_runs = int.from_bytes(p[0:4], "little")
    w = p[4]
    vals = np.frombuffer(p[8:8 + n_runs * w], dtype=dtype)
    lens = np.frombuffer(p[8 + n_runs * w:8 + n_runs * (w + 4)], dtype="<u4")
    pos = 0
    for x, ln in zip(vals, lens):
        ln = min(int(ln), n - pos)
        out[pos:pos + ln] = x
        pos += ln
        if pos == n:
            break
    return out
