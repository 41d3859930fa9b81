"""Nullable GROUP BY keys on the GPU, through the C-ABI, against the oracle:
the LDS direct table (generic and plan-specialised kernels), two keys, ORDER
BY the key both ways with LIMIT, HAVING over the key, radix partitioning,
and a batch."""
import pytest

from helpers import assert_same_rows, rows_of
from test_gpu_parity import REL, check_vs_oracle, float_cols_of
from test_nullable_keys import AGGS, null_key_table
from thindb_amd.exec import (GpuBatchRun, IsNull, Not, Or, Predicate, Request, SortSpec)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("jit", ["0", "1"])
def test_nullable_keys_vs_oracle(session, oracle_lib, monkeypatch, jit):
    monkeypatch.setenv("THINDB_GPU_JIT", jit)
    t = null_key_table()
    for req in (Request(["k1"], AGGS),
                Request(["k2", "k1"], AGGS),
                Request(["k1", "k3"], AGGS, where_filter=Predicate("w", "lt", 30)),
                Request(["k1"], AGGS, [SortSpec("k1")]),
                Request(["k2", "k1"], AGGS, [SortSpec("k1", True), SortSpec("k2")], limit=15, offset=2),
                Request(["k2", "k1"], AGGS, [SortSpec("s", True)], limit=10),
                Request(["k2", "k1"], AGGS, having_filter=Or((IsNull("k1"), Predicate("k2", "eq", 2)))),
                Request(["k2", "k1"], AGGS, having_filter=Not(Predicate("k1", "gt", 0)))):
        got, want, st = check_vs_oracle(session, oracle_lib, t, req)
        if not req.having_filter:
            assert any(None in r[:len(req.group_cols)] for r in got)
    # ORDER BY the key: NULL first ascending, last descending
    got, _, _ = check_vs_oracle(session, oracle_lib, t, Request(["k1"], AGGS, [SortSpec("k1")]))
    assert got[0][0] is None
    got, _, _ = check_vs_oracle(session, oracle_lib, t, Request(["k1"], AGGS, [SortSpec("k1", True)]))
    assert got[-1][0] is None


def test_nullable_keys_radix(session, oracle_lib, monkeypatch):
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    t = null_key_table(seed=9, n=300_000, rg_rows=4099)
    for req in (Request(["k2", "k1", "k3"], AGGS),
                Request(["k2", "k1", "k3"], AGGS, [SortSpec("k1"), SortSpec("k2", True), SortSpec("k3")], 30)):
        got, want, st = check_vs_oracle(session, oracle_lib, t, req)
        assert st.global_mode == 3


def test_nullable_keys_batched(session, oracle_lib):
    t = null_key_table(seed=10)
    reqs = [Request(["k1"], AGGS[:3]), Request(["k2", "k1"], AGGS[:3]), Request(["k1", "k3"], AGGS[:3]),
            Request(["k2", "k1"], AGGS[:3], having_filter=IsNull("k1"))]
    run = GpuBatchRun(session, [(t, r) for r in reqs])
    try:
        for r, b in zip(reqs, run.execute()):
            assert_same_rows(rows_of(b), rows_of(oracle_lib.run(t, r, 2)), float_cols_of(t, r), REL,
                             len(r.group_cols))
    finally:
        run.close()
