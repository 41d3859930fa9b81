"""Encoded column blocks (tdbg_column_stage_encoded): the payload as thinDB's
block cache holds it — validity bitmap first, then a raw, FOR or RLE body —
expanded in HBM, then queried.  Results equal the oracle's over the decoded
table; malformed payloads are refused."""
import numpy as np
import pytest

from helpers import ALL_AGGS, bitmap_bytes, decode_payload, for_payload, random_table, rle_payload, rows_of
from thindb_amd.exec import And, AggSpec, Column, Predicate, Request, SortSpec, Table, try_build_gpu_group
from thindb_amd.gpu import Encoding, InvalidArgument, Type, Unsupported


def test_payload_encoders_round_trip():
    """CPU: the test encoders against a decoder written from the reader."""
    rng = np.random.default_rng(1)
    for dtype, lo, hi in ((np.int32, -50, 50), (np.int64, -2 ** 62, 2 ** 62), (np.int16, -30000, 30000),
                          (np.int8, -3, 4)):
        v = rng.integers(lo, hi, 777).astype(dtype)
        valid = rng.random(777) > 0.2
        for width in (0, 8):
            p = for_payload(v, valid, width)
            got = decode_payload("for", p, 777, dtype, True)
            assert (got[valid] == v[valid]).all()
        p = rle_payload(np.sort(v))
        assert (decode_payload("rle", p, 777, dtype, False) == np.sort(v)).all()


def stage_encoded(session, t, plan):
    """Stages every block of `t` by `plan`: column -> (encoding, FOR width)."""
    for rg in t.row_groups:
        for c in t.schema:
            enc, width = plan.get(c.name, ("raw", 0))
            v = t.columns[c.name][rg.start:rg.start + rg.n_rows]
            valid = t.validity.get(c.name)
            valid = None if valid is None else valid[rg.start:rg.start + rg.n_rows]
            if enc == "for":
                payload, e = for_payload(v, valid, width), Encoding.FOR
            elif enc == "rle":
                payload, e = rle_payload(v, valid), Encoding.RLE
            else:
                payload = (bitmap_bytes(valid) if valid is not None else b"") + np.ascontiguousarray(v).tobytes()
                e = Encoding.RAW
            ci = t.column_index(c.name)
            session.ctx.stage_encoded(t.uid, rg.segment_id, rg.idx, ci, c.type, e, payload, rg.n_rows,
                                      valid is not None)
            session._staged.add((t.uid, rg.segment_id, rg.idx, ci))


@pytest.mark.gpu
def test_encoded_table_vs_oracle(session, oracle_lib):
    from test_gpu_parity import check_vs_oracle
    rng = np.random.default_rng(31)
    t = random_table(rng, 40_000, rg_rows=5_001, nulls=True)
    plan = {"k1": ("for", 0), "k2": ("rle", 0), "b": ("for", 8), "s": ("for", 2), "t": ("for", 0),
            "m": ("rle", 0), "i": ("raw", 0)}
    print(f"[gpu] encoded staging of {t.n_rows} rows: {plan}", flush=True)
    stage_encoded(session, t, plan)
    where = And((Predicate("i", "gt", -2 ** 30), Predicate("s", "neq", 0), Predicate("t", "lt", 8 * 10 ** 11)))
    for req in [Request(["k1"], ALL_AGGS),
                Request(["k1", "k2"], ALL_AGGS, where_filter=where),
                Request(["k2"], ALL_AGGS, [SortSpec("sb", True)], limit=5),
                Request(["k1"], ALL_AGGS, [SortSpec("mint"), SortSpec("k1", True)], limit=17, offset=3)]:
        check_vs_oracle(session, oracle_lib, t, req)
    session.evict_table(t)


@pytest.mark.gpu
def test_rle_runs_clamped_and_short(session):
    """Runs past the block's rows are cut; rows past the last run read 0."""
    n = 1000
    k = np.zeros(n, np.int32)
    v = np.repeat(np.arange(10, dtype=np.int64), 100)  # 10 runs of 100
    t = Table([Column("k", Type.INT), Column("v", Type.BIGINT)], {"k": k, "v": v}, rg_rows=n)
    rg = t.row_groups[0]
    for lengths, want in (([100] * 9 + [90], v.sum() - 9 * 10),            # 990 rows covered: the last 10 read 0
                          ([100] * 9 + [500], v.sum())):                    # the last run clamped to the block
        session.evict_table(t)
        session.ctx.stage_encoded(t.uid, rg.segment_id, rg.idx, 0, Type.INT, Encoding.RAW, k.tobytes(), n)
        session.ctx.stage_encoded(t.uid, rg.segment_id, rg.idx, 1, Type.BIGINT, Encoding.RLE,
                                  rle_payload(v, None, lengths), n)
        session._staged.update({(t.uid, rg.segment_id, rg.idx, 0), (t.uid, rg.segment_id, rg.idx, 1)})
        q = try_build_gpu_group(session, t, Request(["k"], [AggSpec("sum", "v", "s"), AggSpec("count", None, "c")]))
        assert rows_of(q.next()) == [(0, int(want), n)]
        q.deinit()
    session.evict_table(t)


@pytest.mark.gpu
def test_malformed_payloads_refused(session):
    ctx = session.ctx
    v = np.arange(100, dtype=np.int64)
    good = for_payload(v)
    cases = [(Encoding.FOR, good[:-1], InvalidArgument),                        # codes truncated
             (Encoding.FOR, good[:16] + bytes([3, 0, 0, 0]) + good[20:], InvalidArgument),  # width 3
             (Encoding.FOR, (1 << 70).to_bytes(16, "little") + good[16:], InvalidArgument),  # base past 64 bits
             (Encoding.RLE, rle_payload(v.astype(np.int32)), InvalidArgument),  # 4-byte runs for an 8-byte column
             (Encoding.RLE, rle_payload(v)[:-4], InvalidArgument),              # lengths truncated
             (Encoding.DICT, good, Unsupported)]
    for enc, payload, err in cases:
        with pytest.raises(err):
            ctx.stage_encoded(987654, 0, 0, 0, Type.BIGINT, enc, payload, 100)
    with pytest.raises(Unsupported):
        ctx.stage_encoded(987654, 0, 0, 0, Type.DOUBLE, Encoding.FOR, good, 100)
    ctx.drop_table(987654)
