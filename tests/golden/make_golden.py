"""Writes the golden fixtures of the grouped-aggregate path.

Each fixture restates one known-answer test of the reference (thinDB source
tree, read as text — nothing of the reference is imported or run): the
input rows, built by the same rule the test uses, and the expected results
the test asserts.  Expected values that the reference states as constants
are copied as data and re-derived here with exact integer arithmetic.

  wrap_sum.json      tests/integration/aggregate_test.zig:1282-1425
                     (SUM(BIGINT) wraps; WRAP_G1/G2/G3 constants :1298-1301)
  decimal_group.json tests/integration/decimal_test.zig:447-480
                     (grouped SUM/AVG over DECIMAL(12,4))
  float_group.json   tests/integration/v2_group_topn_test.zig:180-196
                     (SUM/MIN/MAX/COUNT/AVG over DOUBLE and FLOAT)
  heavy_groups.json  tests/integration/v2_group_topn_test.zig:2635-2700
                     (heavy group g holds 6000-250g rows among 245,000
                     singleton groups; COUNT DESC LIMIT 10 with SUM(v))
  ranked_pages.json  tests/integration/v2_group_topn_test.zig:1408-1460
                     (140,009 groups; five ORDER BY ... LIMIT/OFFSET pages,
                     ties in the ORDER BY value broken by the group key)
  nullable_aggs.json tests/integration/v2_group_topn_test.zig:1283-1400
                     (20,011 groups, nullable INT/BIGINT/DOUBLE inputs, group
                     0 all NULL; ORDER BY SUM(id) DESC pages)

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import math
from pathlib import Path

HERE = Path(__file__).resolve().parent
MASK64 = (1 << 64) - 1


def wrap64(x: int) -> int:
    x &= MASK64
    return x - (1 << 64) if x >> 63 else x


def wrap_sum() -> dict:
    head = [
        (1, 1, 1, 9_000_000_000_000_000_000),
        (2, 1, 1, 9_000_000_000_000_000_000),
        (3, 1, 1, 9_000_000_000_000_000_000),
        (4, 2, 2, 5_000_000_000_000_000_000),
        (5, 2, 2, 5_000_000_000_000_000_000),
        (6, 3, 3, 1),
        (7, 3, 3, 2),
        (8, 3, 3, None),
    ]
    filler_rows = 70_000  # (id 100+i, g 4, hk 100+i, v 0)
    g1 = wrap64(3 * 9_000_000_000_000_000_000)
    g2 = wrap64(2 * 5_000_000_000_000_000_000)
    g3 = 3
    assert g1 == 8553255926290448384 and g2 == -8446744073709551616
    return {
        "source": "tests/integration/aggregate_test.zig:1282-1425",
        "schema": [["id", "BIGINT", False], ["g", "INT", False], ["hk", "INT", False], ["v", "BIGINT", True]],
        "head": [list(r) for r in head],
        "filler": {"rows": filler_rows, "id0": 100, "g": 4, "hk0": 100, "v": 0},
        "row_group_size": 4096,
        "WRAP_G1": g1, "WRAP_G2": g2, "WRAP_G3": g3,
        "queries": [
            {"sql": "SELECT g, SUM(v) AS s FROM w GROUP BY g ORDER BY g",
             "keys": ["g"], "order": [["g", False]], "limit": 0,
             "expect": [[1, g1], [2, g2], [3, g3], [4, 0]]},
            {"sql": "SELECT g, SUM(v) AS s FROM w GROUP BY g ORDER BY s DESC LIMIT 2",
             "keys": ["g"], "order": [["s", True]], "limit": 2,
             "expect": [[1, g1], [3, g3]]},
            {"sql": "SELECT hk, SUM(v) AS s FROM w GROUP BY hk ORDER BY s DESC LIMIT 2",
             "keys": ["hk"], "order": [["s", True]], "limit": 2,
             "expect": [[1, g1], [3, g3]]},
            {"sql": "SELECT hk, SUM(v) AS s FROM w GROUP BY hk ORDER BY hk LIMIT 3",
             "keys": ["hk"], "order": [["hk", False]], "limit": 3,
             "expect": [[1, g1], [2, g2], [3, g3]]},
            # HAVING on the wrapped value (:1410-1412, :1438-1440)
            {"sql": "SELECT g, SUM(v) AS s FROM w GROUP BY g HAVING SUM(v) < 0",
             "keys": ["g"], "order": [], "limit": 0, "having": ["s", "lt", 0],
             "expect": [[2, g2]]},
            {"sql": "SELECT hk, SUM(v) AS s FROM w GROUP BY hk HAVING SUM(v) < 0",
             "keys": ["hk"], "order": [], "limit": 0, "having": ["s", "lt", 0],
             "expect": [[2, g2]]},
        ],
        "all_hk_groups": 3 + filler_rows,
    }


def decimal_group() -> dict:
    rows = [(1, 1, 15000), (2, 1, 25000), (3, 1, 30000), (4, 2, 100000), (5, 2, 200000)]  # v scaled by 10^4
    sums = {1: 70000, 2: 300000}
    return {
        "source": "tests/integration/decimal_test.zig:447-480",
        "schema": [["id", "INT", False], ["grp", "INT", False], ["v", "DECIMAL64", True, 4]],
        "rows": [list(r) for r in rows],
        "sum": [[1, sums[1]], [2, sums[2]]],
        # the reference asserts |avg - 2.333333| <= 1e-5 and |avg - 15| <= 1e-9
        "avg_approx": [[1, 2.333333, 1e-5], [2, 15.0, 1e-9]],
        # exact closed form f64(sum)/f64(n)/10^4 of v2_lowcard_group.zig:2128-2139
        "avg_exact": [[1, float(70000) / 3.0 / 10000.0], [2, float(300000) / 2.0 / 10000.0]],
    }


def float_group() -> dict:
    rows = [(1, 1, 1.5, 1.5), (2, 1, 2.5, 2.5), (3, 1, 3.0, 3.0), (4, 2, 10.25, 10.25), (5, 2, -4.25, -4.25)]
    return {
        "source": "tests/integration/v2_group_topn_test.zig:180-196",
        "schema": [["id", "BIGINT", False], ["k", "SMALLINT", False], ["sd", "DOUBLE", False], ["sf", "FLOAT", False]],
        "rows": [list(r) for r in rows],
        "expect": {"1": {"sum": 7.0, "min": 1.5, "max": 3.0, "count": 3, "avg": 7.0 / 3.0},
                   "2": {"sum": 6.0, "min": -4.25, "max": 10.25, "count": 2, "avg": 3.0}},
    }


def heavy_groups() -> dict:
    heavy, unique_rows = 12, 245_000
    counts, sums = [0] * heavy, [0] * heavy
    i = 0
    for g in range(heavy):
        for _ in range(6000 - 250 * g):
            counts[g] += 1
            sums[g] += i % 7
            i += 1
    total = i + unique_rows
    return {
        "source": "tests/integration/v2_group_topn_test.zig:2635-2700",
        "rule": "row i (insertion order): heavy rows first, group g gets 6000-250g rows with k=g, v=i%7; "
                "then 245000 rows k=1000+j, v=i%7; physical order sorts rows by id=(i*7919)%total",
        "heavy_groups": heavy, "unique_rows": unique_rows, "total": total, "row_group_size": 1024,
        "top10": [[g, counts[g], sums[g]] for g in range(10)],
    }


def ranked_rows(group_count: int = 140_009):
    """Rows of the ranked-emit test: group k (key g = k * 1,000,003) holds
    1 + k % 3 rows, row j of it v = (k * 31 + j * 7) % 17, ids in insertion
    order."""
    out = []
    for k in range(group_count):
        for j in range(1 + k % 3):
            out.append((len(out), k * 1_000_003, (k * 31 + j * 7) % 17))
    return out


def ranked_pages() -> dict:
    """Expected pages from a plain sort of the groups.  The reference's
    result order is the ORDER BY items, then the group key ascending
    (rowOrderBefore, src/exec/v2_shape_group_topn.zig:1327-1335; the test's
    comment: ties "break on the group key"), so every page is determined."""
    group_count = 140_009
    grp = {}
    for _, g, v in ranked_rows(group_count):
        e = grp.setdefault(g, [0, 0, None])  # count, sum, min
        e[0] += 1
        e[1] += v
        e[2] = v if e[2] is None else min(e[2], v)
    G = list(grp.items())
    cases = []
    # (sql, output name, value of a group, sort key of a group, limit, offset, having)
    specs = [
        ("SELECT g, MIN(v) AS m FROM rank_rows GROUP BY g ORDER BY m DESC", "min",
         lambda e: e[2], lambda g, e: (-e[2], g), 10, 3, None),
        ("SELECT g, SUM(v) AS s FROM rank_rows GROUP BY g HAVING COUNT(*) > 1 ORDER BY s", "sum",
         lambda e: e[1], lambda g, e: (e[1], g), 20, 0, "count>1"),
        ("SELECT g, COUNT(*) AS c FROM rank_rows GROUP BY g ORDER BY c, g DESC", "count",
         lambda e: e[0], lambda g, e: (e[0], -g), 7, 11, None),
        ("SELECT g, AVG(v) AS a FROM rank_rows GROUP BY g ORDER BY a DESC, g", "avg",
         lambda e: float(e[1]) / float(e[0]), lambda g, e: (-(float(e[1]) / float(e[0])), g), 5, 0, None),
        ("SELECT g, SUM(v) AS s FROM rank_rows GROUP BY g ORDER BY s DESC, g", "sum",
         lambda e: e[1], lambda g, e: (-e[1], g), 40, 39_990, None),
    ]
    for sql, func, val, key, limit, offset, having in specs:
        rows = [(g, e) for g, e in G if having is None or e[0] > 1]
        rows.sort(key=lambda ge: key(*ge))
        page = [[g, val(e)] for g, e in rows[offset:offset + limit]]
        assert len(page) == limit
        cases.append({"sql": f"{sql} LIMIT {limit} OFFSET {offset}", "func": func, "limit": limit,
                      "offset": offset, "having_count_gt_1": having is not None, "page": page})
    return {
        "source": "tests/integration/v2_group_topn_test.zig:1408-1460",
        "rule": "group k in 0..140008: g = k * 1000003 (BIGINT), 1 + k % 3 rows, row j: "
                "v = (k * 31 + j * 7) % 17 (INT), id = insertion index (BIGINT)",
        "group_count": group_count, "row_group_size": 8192, "cases": cases,
    }


def nullable_rows(group_count: int = 20_011):
    """Rows of the nullable-aggregate test: row i, g = (i // 2) % group_count
    (a nullable INT column with no NULL), v/b/d NULL in group 0 and on rows
    i % 7 == 0 / i % 5 == 0 / i % 3 == 0."""
    out = []
    for i in range(group_count * 6):
        g = (i // 2) % group_count
        z = g == 0
        v = None if (z or i % 7 == 0) else (i * 37) % 1000 - 500
        b = None if (z or i % 5 == 0) else i * 3_000_000_019
        d = None if (z or i % 3 == 0) else (float((i * 13) % 512) - 256.0) * 0.25
        out.append((i, g, v, b, d))
    return out


def nullable_aggs() -> dict:
    """The test's expected values per group (sums exact: d is a multiple of
    0.25 far below 2^53), ranked by SUM(id) DESC (tie-free: 12g + const)."""
    group_count = 20_011
    E = [dict(id_sum=0, count=0, v=[], b=[], d=[]) for _ in range(group_count)]
    for i, g, v, b, d in nullable_rows(group_count):
        e = E[g]
        e["id_sum"] += i
        e["count"] += 1
        for name, x in (("v", v), ("b", b), ("d", d)):
            if x is not None:
                e[name].append(x)

    def int_row(g):
        e = E[g]
        v = e["v"]
        return [g, e["id_sum"], e["count"], sum(v) if v else None, float(sum(v)) / float(len(v)) if v else None,
                min(v) if v else None, max(v) if v else None, sum(e["b"]) if e["b"] else None]

    def float_row(g):
        e = E[g]
        b, d = e["b"], e["d"]
        return [g, e["id_sum"], float(sum(b)) / float(len(b)) if b else None, math.fsum(d) if d else None,
                min(d) if d else None, max(d) if d else None]

    ranked = list(range(group_count - 1, -1, -1))  # SUM(id) DESC
    ids = [E[g]["id_sum"] for g in ranked]
    assert all(a > b for a, b in zip(ids, ids[1:]))
    out = {"source": "tests/integration/v2_group_topn_test.zig:1283-1400",
           "rule": "row i in 0..6*20011: id = i BIGINT, g = (i // 2) % 20011 INT nullable, "
                   "v INT nullable = NULL if g == 0 or i % 7 == 0 else (i * 37) % 1000 - 500, "
                   "b BIGINT nullable = NULL if g == 0 or i % 5 == 0 else i * 3000000019, "
                   "d DOUBLE nullable = NULL if g == 0 or i % 3 == 0 else ((i * 13) % 512 - 256) * 0.25",
           "group_count": group_count, "row_group_size": 4096}
    for name, fn in (("int", int_row), ("float", float_row)):
        rows = [fn(g) for g in ranked]
        out[name] = {"page_25_7": rows[7:32],
                     "digest_all": hashlib.sha256(json.dumps(rows).encode()).hexdigest()}
    return out


def main() -> None:
    for name, fn in [("wrap_sum", wrap_sum), ("decimal_group", decimal_group), ("float_group", float_group),
                     ("heavy_groups", heavy_groups), ("ranked_pages", ranked_pages),
                     ("nullable_aggs", nullable_aggs)]:
        (HERE / f"{name}.json").write_text(json.dumps(fn(), indent=1) + "\n")
        print("wrote", name)


if __name__ == "__main__":
    main()
