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

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
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


def main() -> None:
    for name, fn in [("wrap_sum", wrap_sum), ("decimal_group", decimal_group), ("float_group", float_group),
                     ("heavy_groups", heavy_groups)]:
        (HERE / f"{name}.json").write_text(json.dumps(fn(), indent=1) + "\n")
        print("wrote", name)


if __name__ == "__main__":
    main()
