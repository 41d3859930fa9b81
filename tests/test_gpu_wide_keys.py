"""Wide and sparse GROUP BY keys on the GPU, against the oracle: hash
partitions (global_mode 4), where rows are routed by the top bits of their
key's hash and each leaf folds into an LDS hash table.

* two full-range BIGINT keys: a key domain of 2^128, the packed 128-bit key
  (the silo's <=128-bit packing, v2_shape_group_topn.zig:2051-2094; routing
  by hash bits, silo_group_core.zig:5226-5235);
* a full-range BIGINT and a nullable INT of both signs (97 packed bits; ties
  broken field by field);
* a sparse single BIGINT key (domain 2^50, one row per group or so): the
  relative key in one word, which would otherwise take per-row HBM atomics;
* HAVING in the leaf pass, ORDER BY COUNT(*) through the leaves' count
  histogram, a WHERE filter, and the leaf-overflow retry (an est_groups far
  below the real group count).
"""
import numpy as np
import pytest

from helpers import check_with_oracle, rows_of, wide_key_table
from test_gpu_parity import REL, check_vs_oracle, run_gpu
from thindb_amd.exec import AggSpec, And, Column, Not, Or, Predicate, Request, SortSpec, Table
from thindb_amd.gpu import Type

pytestmark = pytest.mark.gpu

AGGS = [AggSpec("count", None, "c"), AggSpec("sum", "v", "sv"), AggSpec("min", "w", "mnw"),
        AggSpec("max", "w", "mxw"), AggSpec("avg", "d", "ad"), AggSpec("sum", "x", "sx"),
        AggSpec("sum", "w", "sw"), AggSpec("count", "x", "cx")]


@pytest.fixture(scope="module")
def big_wide():
    return wide_key_table(n=3_000_000, groups=1_200_000, seed=7)


def test_two_full_range_keys_vs_oracle(session, oracle_lib, big_wide):
    t = big_wide
    got, want, st = check_vs_oracle(session, oracle_lib, t, Request(["a", "b"], AGGS), dop=8)
    print(f"  {st.n_groups} groups, global_mode {st.global_mode}, lds_mode {st.lds_mode}", flush=True)
    assert st.global_mode == 4 and st.n_groups == len(want) == 1_200_000
    for order, limit, offset in [([SortSpec("c", True)], 100, 0), ([SortSpec("b", True), SortSpec("a")], 50, 7),
                                 ([SortSpec("sx")], 80, 0), ([SortSpec("mxw", True)], 1000, 0)]:
        _, _, st = check_vs_oracle(session, oracle_lib, t, Request(["a", "b"], AGGS, order, limit, offset), dop=8)
        assert st.global_mode == 4


def test_wide_key_with_nullable_signed_part(session, oracle_lib):
    t = wide_key_table(n=1_500_000, groups=400_000, seed=3, nullable=True)
    for keys, order, limit in [(["a", "c"], [], 0), (["a", "c"], [SortSpec("c", True)], 60),
                               (["c", "a"], [SortSpec("sw", True)], 40), (["a", "c"], [SortSpec("c")], 0)]:
        _, want, st = check_vs_oracle(session, oracle_lib, t, Request(keys, AGGS, order, limit), dop=8)
        assert st.global_mode == 4


def test_wide_keys_having_and_where(session, oracle_lib, big_wide):
    t = big_wide
    h = Or((Predicate("c", "gte", 3), And((Predicate("sv", "lt", 0), Predicate("a", "gt", 0)))))
    w = Or((Predicate("v", "gt", 0), Predicate("x", "lt", -100)))
    _, want, st = check_vs_oracle(session, oracle_lib, t, Request(["a", "b"], AGGS, having_filter=h), dop=8)
    assert st.global_mode == 4 and 0 < len(want) < 1_200_000
    check_vs_oracle(session, oracle_lib, t, Request(["a", "b"], AGGS, [SortSpec("sv")], 25, having_filter=Not(h),
                                                    where_filter=w), dop=8)


def test_sparse_bigint_key_takes_hash_partitions(session, oracle_lib):
    rng = np.random.default_rng(11)
    n = 2_000_000
    k = rng.integers(-2 ** 49, 2 ** 49, n)
    k[: n // 4] = k[n // 2: n // 2 + n // 4]  # a quarter of the rows repeat other keys
    t = Table([Column("k", Type.BIGINT), Column("v", Type.INT), Column("d", Type.DOUBLE)],
              {"k": k, "v": rng.integers(-1000, 1000, n).astype(np.int32), "d": rng.random(n)})
    aggs = [AggSpec("count", None, "c"), AggSpec("sum", "v", "s"), AggSpec("avg", "d", "a"), AggSpec("min", "v", "m")]
    _, want, st = check_vs_oracle(session, oracle_lib, t, Request(["k"], aggs), dop=8)
    assert st.global_mode == 4 and st.n_groups == len(want)
    _, _, st = check_vs_oracle(session, oracle_lib, t, Request(["k"], aggs, [SortSpec("c", True)], 10), dop=8)
    assert st.global_mode == 4


def test_forced_hash_partitions_small_tables(session, oracle_lib, monkeypatch):
    """THINDB_GPU_RADIX_HASH=1 routes ordinary shapes (dense and narrow keys,
    one and two scatter levels) through hash partitions."""
    monkeypatch.setenv("THINDB_GPU_RADIX_HASH", "1")
    rng = np.random.default_rng(2)
    n = 400_000
    t = Table([Column("s", Type.SMALLINT), Column("i", Type.INT, True), Column("v", Type.BIGINT),
               Column("f", Type.FLOAT)],
              {"s": rng.integers(-300, 300, n).astype(np.int16), "i": rng.integers(-5000, 5000, n).astype(np.int32),
               "v": rng.integers(-10 ** 12, 10 ** 12, n), "f": rng.random(n).astype(np.float32)},
              {"i": rng.random(n) > 0.1})
    aggs = [AggSpec("count", None, "c"), AggSpec("sum", "v", "sv"), AggSpec("max", "f", "mf"),
            AggSpec("count", "i", "ci")]
    for keys, order, limit in [(["s"], [], 0), (["s", "i"], [SortSpec("sv", True)], 30), (["i", "s"], [SortSpec("c")], 0)]:
        _, _, st = check_vs_oracle(session, oracle_lib, t, Request(keys, aggs, order, limit,
                                                                   where_filter=Predicate("v", "gt", -9 * 10 ** 11)))
        assert st.global_mode == 4


def test_leaf_overflow_retries_with_more_leaves(session, oracle_lib, big_wide):
    """An est_groups 1000x below the real count sizes the leaves' tables for
    too few groups: the tables overflow, the plan partitions again into more
    leaves, and the result is whole."""
    t = big_wide
    req = Request(["a", "b"], AGGS[:4], [SortSpec("c", True)], 20, est_groups=1200)
    got, want, st = check_vs_oracle(session, oracle_lib, t, req, dop=8)
    assert st.global_mode == 4 and st.n_groups == 1_200_000


@pytest.mark.parametrize("env", [{"THINDB_GPU_RADIX_HASH_TAG_BITS": "2"}, {"THINDB_GPU_RADIX_HASH_EXACT": "1"}])
def test_hashed_fold_fallbacks(session, oracle_lib, monkeypatch, env):
    """The leaf fold's exact form: reached through a tag collision (tags
    narrowed to two bits, so rows meet other keys' tags on their probe paths
    and their leaves are folded again), and run alone."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    t = wide_key_table(n=600_000, groups=150_000, seed=5)
    _, want, st = check_vs_oracle(session, oracle_lib, t, Request(["a", "b"], AGGS), dop=8)
    assert st.global_mode == 4 and st.n_groups == len(want) == 150_000
    check_vs_oracle(session, oracle_lib, t, Request(["a", "b"], AGGS, [SortSpec("c", True)], 30), dop=8)
