"""WHERE trees on the GPU, through the C-ABI, against the oracle: the
generic and plan-specialised scan kernels, the radix partition passes, a
batch, COUNT(DISTINCT) and HAVING on top, and malformed trees rejected."""
import numpy as np
import pytest

from helpers import ALL_AGGS, assert_same_rows, random_table, rows_of
from test_gpu_parity import REL, check_vs_oracle, float_cols_of
from test_where_tree import TREES
from thindb_amd import gpu
from thindb_amd.exec import (AggSpec, And, GpuBatchRun, IsNull, Not, Or, Predicate, Request, SortSpec)
from thindb_amd.gpu import HavingKind, PredOp, Type

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("jit", ["0", "1"])
def test_where_trees_vs_oracle(session, oracle_lib, monkeypatch, jit):
    monkeypatch.setenv("THINDB_GPU_JIT", jit)
    rng = np.random.default_rng(31)
    t = random_table(rng, 40_000, rg_rows=701, nulls=True)
    for tree in TREES:
        print(f"[where] {tree}", flush=True)
        _, _, st = check_vs_oracle(session, oracle_lib, t, Request(["k1", "k2"], ALL_AGGS, where_filter=tree))
        check_vs_oracle(session, oracle_lib, t, Request(["k2"], ALL_AGGS, [SortSpec("sb", True)], 3,
                                                        where_filter=Not(tree)))
        if jit == "1" and len({a.col for a in ALL_AGGS} | {"k1", "k2"}) <= 16:
            assert st.specialized in (0, 1)


def test_where_trees_radix_distinct_having(session, oracle_lib, monkeypatch):
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    rng = np.random.default_rng(32)
    t = random_table(rng, 60_000, rg_rows=4099, nulls=True)
    aggs = ALL_AGGS + [AggSpec("count_distinct", "k2", "dk")]
    for tree in TREES[:4]:
        _, _, st = check_vs_oracle(session, oracle_lib, t, Request(["k1", "k2"], aggs, where_filter=tree,
                                                                   est_groups=10,
                                                                   having_filter=Predicate("c", "gt", 20)))
        assert st.global_mode == 3


def test_where_trees_batched(session, oracle_lib):
    rng = np.random.default_rng(33)
    t = random_table(rng, 30_000, rg_rows=1000, nulls=True)
    reqs = [Request(["k1"], ALL_AGGS, where_filter=tree) for tree in TREES]
    run = GpuBatchRun(session, [(t, r) for r in reqs])
    try:
        for r, b in zip(reqs, run.execute()):
            assert_same_rows(rows_of(b), rows_of(oracle_lib.run(t, r, 2)), float_cols_of(t, r), REL, 1)
    finally:
        run.close()


def test_where_tree_rejects_malformed(session):
    rng = np.random.default_rng(34)
    t = random_table(rng, 2000)
    session.stage_table(t, ["k1", "i"])
    keys = [(t.column_index("k1"), Type.INT)]
    aggs = [(gpu.AggOp.COUNT_STAR, -1, Type.BIGINT, 0, False)]
    leaf = (HavingKind.CMP, t.column_index("i"), Type.INT, PredOp.GT, 0)
    for nodes in ([(HavingKind.NOT,)], [leaf, leaf], [leaf, (HavingKind.OR, 3)],
                  [(HavingKind.CMP, t.column_index("i"), Type.INT, 9, 0)]):
        with pytest.raises(gpu.InvalidArgument):
            gpu.Plan(session.ctx, t.uid, keys, aggs, where=nodes)
    with pytest.raises(gpu.Unsupported):
        gpu.Plan(session.ctx, t.uid, keys, aggs, where=[leaf] * 13 + [(HavingKind.AND, 13)])
