"""Random grouped queries on the GPU, through the C-ABI, against the oracle:
one to three keys (nullable and wide ones among them), random aggregate
lists (COUNT(DISTINCT) included), WHERE conjunctions and trees, HAVING,
ORDER BY over outputs with OFFSET/LIMIT, over typed tables with NULLs.  The
shapes the operator declines are counted and skipped (a handful at most)."""
import numpy as np
import pytest

from helpers import random_table
from test_gpu_parity import check_vs_oracle
from thindb_amd.exec import (AggSpec, And, IsNotNull, IsNull, Not, Or, Predicate, Request, SortSpec,
                             try_build_gpu_group)

pytestmark = pytest.mark.gpu

KEY_SETS = [["k1"], ["k2"], ["s"], ["i"], ["t"], ["k1", "k2"], ["k2", "s"], ["k1", "s"], ["s", "k1", "k2"],
            ["k2", "i"], ["k1", "k2", "i"]]
AGGS = [("count", None), ("count", "s"), ("sum", "b"), ("sum", "i"), ("avg", "i"), ("min", "s"), ("max", "b"),
        ("sum", "d"), ("avg", "d"), ("min", "f"), ("max", "d"), ("sum", "m"), ("avg", "m"), ("min", "t"),
        ("max", "k2"), ("count_distinct", "k2"), ("count_distinct", "s"), ("count_distinct", "f")]
FILTER_COLS = ["k1", "k2", "b", "i", "s", "d", "f", "m", "t"]
OPS = ["eq", "neq", "lt", "lte", "gt", "gte"]


def _leaf(rng, t):
    c = FILTER_COLS[rng.integers(len(FILTER_COLS))]
    v = t.columns[c][rng.integers(t.n_rows)]
    lit = float(v) if t.column(c).type.name in ("FLOAT", "DOUBLE") else int(v)
    return Predicate(c, OPS[rng.integers(len(OPS))], lit)


def _where(rng, t):
    k = rng.integers(7)
    if k == 0:
        return None
    if k == 1:
        return _leaf(rng, t)
    if k == 2:
        return And((_leaf(rng, t), _leaf(rng, t)))
    if k == 3:
        return Or((_leaf(rng, t), _leaf(rng, t)))
    if k == 4:
        return Not(_leaf(rng, t))
    if k == 5:
        return And((_leaf(rng, t), Or((IsNull("s"), Not(_leaf(rng, t))))))
    return And((IsNotNull("b"), _leaf(rng, t)))


def _request(rng, t):
    keys = KEY_SETS[rng.integers(len(KEY_SETS))]
    picks = rng.choice(len(AGGS), size=int(rng.integers(1, 6)), replace=False)
    aggs = [AggSpec(AGGS[i][0], AGGS[i][1], f"a{j}") for j, i in enumerate(picks)]
    outputs = [a.as_ for a in aggs]
    order = []
    for _ in range(int(rng.integers(0, 3))):
        order.append(SortSpec(outputs[rng.integers(len(outputs))], bool(rng.integers(2))))
    if order or rng.integers(2):
        order += [SortSpec(k, bool(rng.integers(2))) for k in keys]  # ties broken by every key
    limit = int(rng.choice([0, 5, 50])) if order else 0
    offset = int(rng.choice([0, 3])) if limit else 0
    having = None
    counts = [a.as_ for a in aggs if a.func.startswith("count")]
    if counts and rng.integers(3) == 0:
        having = Predicate(counts[0], "gte", int(rng.integers(1, 4)))
    return Request(keys, aggs, order, limit, offset, where_filter=_where(rng, t), having_filter=having)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_queries_vs_oracle(session, oracle_lib, seed):
    rng = np.random.default_rng(100 + seed)
    t = random_table(rng, int(rng.integers(5_000, 30_000)), rg_rows=int(rng.choice([512, 1000, 4096])), nulls=True)
    declined, ran = 0, 0
    for case in range(30):
        req = _request(rng, t)
        q = try_build_gpu_group(session, t, req)
        if q is None:
            declined += 1
            continue
        q.deinit()
        print(f"[random] seed {seed} case {case}: {req}", flush=True)
        check_vs_oracle(session, oracle_lib, t, req)
        ran += 1
    assert declined <= 3 and ran >= 27, (declined, ran)


@pytest.mark.parametrize("seed", [4, 5])
def test_random_batch_vs_oracle(session, oracle_lib, seed):
    """24 random queries (no LIMIT / OFFSET / COUNT(DISTINCT), which a batch
    declines) in one batched launch, each against the oracle."""
    from dataclasses import replace

    from helpers import assert_same_rows, rows_of
    from test_gpu_parity import REL, float_cols_of
    from thindb_amd.exec import GpuBatchRun
    rng = np.random.default_rng(200 + seed)
    t = random_table(rng, int(rng.integers(5_000, 30_000)), rg_rows=int(rng.choice([512, 1000, 4096])), nulls=True)
    reqs = []
    while len(reqs) < 24:
        r = _request(rng, t)
        aggs = [a for a in r.aggs if a.func != "count_distinct"] or [AggSpec("count", None, "a0")]
        names = [a.as_ for a in aggs]
        having = r.having_filter if r.having_filter is not None and r.having_filter.col in names else None
        reqs.append(replace(r, aggs=aggs, order_specs=(), limit=0, offset=0, having_filter=having))
    print(f"[random batch] seed {seed}: {len(reqs)} queries", flush=True)
    run = GpuBatchRun(session, [(t, r) for r in reqs])
    try:
        for r, b in zip(reqs, run.execute()):
            want = rows_of(oracle_lib.run(t, r, 4))
            assert_same_rows(rows_of(b), want, float_cols_of(t, r), REL, len(r.group_cols))
    finally:
        run.close()


def test_random_radix_vs_oracle(session, oracle_lib, monkeypatch):
    """Random queries over high-cardinality keys through radix partitioning
    (compacted and dense leaf output)."""
    from dataclasses import replace
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    rng = np.random.default_rng(300)
    t = random_table(rng, 40_000, rg_rows=4096, nulls=True)
    ran = 0
    for case in range(16):
        keys = [["i"], ["k1", "i"], ["t"], ["i", "k2"], ["s", "i"]][case % 5]
        r = replace(_request(rng, t), group_cols=keys, est_groups=int(rng.choice([0, 100_000_000])))
        r = replace(r, order_specs=tuple(s for s in r.order_specs if s.col not in ("k1", "k2", "s", "i", "t"))
                    + tuple(SortSpec(k, False) for k in keys) if r.order_specs else ())
        q = try_build_gpu_group(session, t, r)
        if q is None:
            continue
        q.deinit()
        print(f"[random radix] case {case}: {r}", flush=True)
        check_vs_oracle(session, oracle_lib, t, r)
        ran += 1
    assert ran >= 14
