"""COUNT(DISTINCT col) on the GPU, through the C-ABI, against the oracle:
the LDS direct table, two keys with filters, ORDER BY / LIMIT / HAVING over
the count, radix partitioning with compacted and dense output, the
high-cardinality workload, float and full-range BIGINT values, and the
batch that declines."""
import numpy as np
import pytest

from helpers import brute_force, rows_of
from test_count_distinct import AGGS, FLOAT_AGGS, _table, float_table
from test_gpu_parity import check_vs_oracle, run_gpu
from thindb_amd import gen, gpu
from thindb_amd.exec import AggSpec, And, Column, GpuBatchRun, Predicate, Request, SortSpec, Table
from thindb_amd.gpu import Type

pytestmark = pytest.mark.gpu


def test_count_distinct_vs_oracle(session, oracle_lib):
    t = _table()
    where = And((Predicate("i", "gt", -2 ** 30), Predicate("d", "gte", -500.0)))
    for req in (Request(["k1"], AGGS),
                Request(["k1", "k2"], AGGS, where_filter=where),
                Request(["k2"], AGGS, [SortSpec("ds", True)], where_filter=where),
                Request(["k1", "k2"], AGGS, [SortSpec("dm", True), SortSpec("k1"), SortSpec("k2")], 20, 3),
                Request(["k1"], AGGS, having_filter=Predicate("ds", "lt", 60)),
                Request(["k1"], AGGS, where_filter=Predicate("i", "gt", 2 ** 31))):
        got, want, _ = check_vs_oracle(session, oracle_lib, t, req)
    # brute force on the plain grouping, and the all-NULL groups
    b, _ = run_gpu(session, t, Request(["k1"], AGGS))
    bf = brute_force(t, Request(["k1"], AGGS))
    for r in rows_of(b):
        assert list(r[1:]) == bf[(r[0],)]
    assert all(r[3] == 0 for r in rows_of(b) if r[0] < -45)


@pytest.mark.parametrize("est", [0, 100_000_000])
def test_count_distinct_radix(session, oracle_lib, monkeypatch, est):
    """Radix partitioning (compacted leaves, then dense output) with the
    distinct count of a column that is not a key."""
    t = gen.table("highcard", 1_200_000, seed=21)
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    aggs = [AggSpec("count", None, "c"), AggSpec("count_distinct", "v", "dv"), AggSpec("sum", "v", "s")]
    got, want, st = check_vs_oracle(session, oracle_lib, t, Request(["user_id", "bucket"], aggs, est_groups=est),
                                    dop=8)
    assert st.global_mode == 3
    check_vs_oracle(session, oracle_lib, t, Request(["user_id", "bucket"], aggs, [SortSpec("dv", True)], 10,
                                                    est_groups=est, having_filter=Predicate("c", "gte", 2)), dop=8)
    # HAVING over the distinct count itself (filled after compaction)
    check_vs_oracle(session, oracle_lib, t, Request(["user_id", "bucket"], aggs, est_groups=est,
                                                    having_filter=Predicate("dv", "gte", 2)), dop=8)


def test_count_distinct_highcard_values(session, oracle_lib):
    """Few groups, many distinct values: GROUP BY bucket,
    COUNT(DISTINCT user_id) (the distinct set holds ~1.2M composites)."""
    t = gen.table("highcard", 1_500_000, seed=22)
    req = Request(["bucket"], [AggSpec("count", None, "c"), AggSpec("count_distinct", "user_id", "u")],
                  [SortSpec("u", True)], 10)
    got, want, _ = check_vs_oracle(session, oracle_lib, t, req, dop=8)
    assert len(got) == 10 and all(r[2] <= r[1] for r in got)


def test_count_distinct_sparse_domain(session, oracle_lib):
    """A 2^41-key sparse domain (the HBM hash table): composites use the
    record index and counts go straight into the records."""
    rng = np.random.default_rng(1)
    n = 200_000
    keys = rng.integers(-2 ** 40, 2 ** 40, 3000)
    wide = Table([Column("k", Type.BIGINT), Column("v", Type.INT, True), Column("w", Type.BIGINT)],
                 {"k": keys[rng.integers(0, 3000, n)], "v": rng.integers(0, 90, n).astype(np.int32),
                  "w": rng.integers(-2 ** 40, 2 ** 40, n)}, {"v": rng.random(n) > 0.3})
    aggs = [AggSpec("count", None, "c"), AggSpec("count_distinct", "v", "dv"), AggSpec("count_distinct", "w", "dw")]
    got, want, st = check_vs_oracle(session, oracle_lib, wide, Request(["k"], aggs))
    assert st.global_mode == 2 and len(got) == 3000
    check_vs_oracle(session, oracle_lib, wide, Request(["k"], aggs, [SortSpec("dv", True), SortSpec("k")], 20,
                                                       having_filter=Predicate("dw", "gt", 60)))


def test_count_distinct_floats_and_wide_values(session, oracle_lib, monkeypatch):
    """The wide form: DOUBLE and FLOAT values (one +0.0, one NaN) and BIGINT
    values across the whole int64 range, over the LDS table, the HBM hash
    table (a sparse second key) and radix partitioning."""
    t = float_table()
    for req in (Request(["k1"], FLOAT_AGGS),
                Request(["k1"], FLOAT_AGGS, [SortSpec("dd", True), SortSpec("k1")], 7, 2,
                        where_filter=Predicate("w", "gt", 0)),
                Request(["g"], FLOAT_AGGS, having_filter=Predicate("df", "gte", 2))):
        got, want, st = check_vs_oracle(session, oracle_lib, t, req)
    assert st.global_mode == 2
    b, _ = run_gpu(session, t, Request(["k1"], FLOAT_AGGS))
    bf = brute_force(t, Request(["k1"], FLOAT_AGGS))
    for r in rows_of(b):
        assert list(r[1:]) == bf[(r[0],)]
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    h = gen.table("highcard", 600_000, seed=23)
    aggs = [AggSpec("count", None, "c"), AggSpec("count_distinct", "w", "dw")]
    for est in (0, 100_000_000):
        got, want, st = check_vs_oracle(session, oracle_lib, h, Request(["user_id", "bucket"], aggs, est_groups=est),
                                        dop=8)
        assert st.global_mode == 3


def test_count_distinct_declines(session):
    """A batch refuses a COUNT(DISTINCT) plan."""
    t = _table()
    with pytest.raises(Exception):
        GpuBatchRun(session, [(t, Request(["k1"], AGGS))])
