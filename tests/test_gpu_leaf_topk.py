"""Per-leaf top-k of the radix path (ORDER BY COUNT(*) ... LIMIT): each
compacted leaf writes only its best OFFSET + LIMIT groups by (row count,
key).  Results against the oracle, and the same groups in the same order as
when the leaves write every group (THINDB_GPU_LEAF_TOPK=0); the reported
group count is every group's, not the candidates'.

Cases: descending and ascending counts with heavy ties at the cut (most
groups hold one to three rows), OFFSET, the largest OFFSET + LIMIT the leaves
cut at (1024) and one past it (every group written), and a cut in the last,
unresolved count bin (a leaf whose groups hold 255 rows or more); key parts of
both signs (ties by the packed key: no cut)."""
import numpy as np
import pytest

from helpers import ALL_AGGS, assert_same_rows, float_cols_of, random_table, rows_of
from test_gpu_parity import REL, check_vs_oracle, run_gpu
from thindb_amd.exec import Request, SortSpec, Table

pytestmark = pytest.mark.gpu


def _table(seed: int, heavy: bool, signed: bool = False) -> Table:
    # (k1, k2) spans 400,000 x 7 keys, over twice the groups: the leaves
    # emit compacted records.  Keys of one sign break ties by the relative
    # key (slot order within a leaf: the leaves cut); k2 in [-3, 3] breaks
    # them part by part instead (every group written)
    rng = np.random.default_rng(seed)
    n = 700_000
    base = random_table(rng, n, rg_rows=65_536, nulls=True)
    cols = dict(base.columns)
    k1 = rng.integers(0, 400_000, n).astype(np.int32)
    k2 = cols["k2"].copy() if signed else (cols["k2"] + 3).astype(np.int8)
    if heavy:  # 100 groups of ~2,100 rows each, in the lowest keys (one leaf)
        h = rng.random(n) < 0.3
        k1[h] = rng.integers(0, 100, int(h.sum())).astype(np.int32)
        k2[h] = 0
    k1[0], k1[1] = 0, 399_999  # the full key span either way
    cols["k1"], cols["k2"] = k1, k2
    return Table(base.schema, cols, base.validity, rg_rows=65_536)


def _same_as_every_group(session, t, req, got, st, monkeypatch):
    monkeypatch.setenv("THINDB_GPU_LEAF_TOPK", "0")
    b, st_all = run_gpu(session, t, req)
    monkeypatch.delenv("THINDB_GPU_LEAF_TOPK")
    every = rows_of(b)
    nk = len(req.group_cols)
    assert [r[:nk] for r in every] == [r[:nk] for r in got]
    # DOUBLE sums differ in the last place between runs (the order of float atomics)
    assert_same_rows(every, got, float_cols_of(t, req), REL, nk)
    assert st.n_groups == st_all.n_groups > 65_536


@pytest.mark.parametrize("heavy", [False, True])
def test_leaf_topk_count_order(session, oracle_lib, monkeypatch, heavy):
    t = _table(31 + heavy, heavy)
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    reqs = [Request(["k1", "k2"], ALL_AGGS, [SortSpec("c", True)], limit=10),
            Request(["k1", "k2"], ALL_AGGS, [SortSpec("c")], limit=37, offset=5),
            Request(["k1", "k2"], ALL_AGGS[:6], [SortSpec("c", True)], limit=1000, offset=24),
            Request(["k1", "k2"], ALL_AGGS[:6], [SortSpec("c")], limit=1025)]
    for req in reqs:
        got, want, st = check_vs_oracle(session, oracle_lib, t, req)
        assert st.global_mode == 3
        _same_as_every_group(session, t, req, got, st, monkeypatch)


def test_leaf_topk_signed_keys(session, oracle_lib, monkeypatch):
    t = _table(33, False, signed=True)
    monkeypatch.setenv("THINDB_GPU_RADIX", "1")
    req = Request(["k1", "k2"], ALL_AGGS[:6], [SortSpec("c", True)], limit=20)
    got, want, st = check_vs_oracle(session, oracle_lib, t, req)
    assert st.global_mode == 3
    _same_as_every_group(session, t, req, got, st, monkeypatch)
