"""Segment files staged straight into HBM, then queried, against the oracle
over the same rows: the typed table is written as thinDB segment files
(several segments of several row groups; raw / FOR / RLE bodies, nullable
columns, no / zstd / LZ4 compression, int_streams forms cycling over the
row groups), each block read on the host and expanded on the GPU through
tdbg_segment_stage, and the grouped queries run over what was staged."""
import numpy as np
import pytest

from helpers import ALL_AGGS, random_table
from segment_file import BlockSpec, SegmentColumn, write_segment
from test_gpu_parity import check_vs_oracle
from thindb_amd import gpu
from thindb_amd.exec import AggSpec, Predicate, Request, SortSpec, Table
from thindb_amd.gpu import Type

pytestmark = pytest.mark.gpu

INT_SPECS = [BlockSpec("for", "zstd", "frame"), BlockSpec("raw", "lz4", "delta", True), BlockSpec("rle", "zstd"),
             BlockSpec("for", "none"), BlockSpec("raw", "zstd", "plain", True), BlockSpec("rle", "lz4", "frame", True)]
FLOAT_SPECS = [BlockSpec("raw", "zstd"), BlockSpec("raw", "none"), BlockSpec("raw", "lz4")]


def _segments_of(t: Table):
    """The table's segments as file images, one per segment id."""
    out = {}
    for seg_id in sorted({rg.segment_id for rg in t.row_groups}):
        rgs = [rg for rg in t.row_groups if rg.segment_id == seg_id]
        r0, r1 = rgs[0].start, rgs[-1].start + rgs[-1].n_rows
        cols = []
        for ci, c in enumerate(t.schema):
            specs = FLOAT_SPECS if c.type in (Type.FLOAT, Type.DOUBLE) else INT_SPECS[ci % 2:] + INT_SPECS[:ci % 2]
            val = t.validity.get(c.name)
            cols.append(SegmentColumn(c.type, t.columns[c.name][r0:r1], None if val is None else val[r0:r1],
                                      c.nullable, specs))
        out[seg_id] = write_segment(cols, seg_id, t.rg_rows)
    return out


def test_segment_files_vs_oracle(session, oracle_lib):
    rng = np.random.default_rng(11)
    t = random_table(rng, 20_000, rg_rows=1024, nulls=True)
    t = Table(t.schema, t.columns, t.validity, rg_rows=1024, segment_rows=6144)
    t.columns["k2"] = np.repeat(rng.integers(-3, 4, 20_000 // 64 + 1), 64)[:20_000].astype(np.int8)  # runs for RLE
    t._packed.clear()
    staged = 0
    for seg_id, img in _segments_of(t).items():
        seg = gpu.Segment(img, len(t.schema))
        assert seg.info().segment_id == seg_id
        print(f"[segment] {seg_id}: {len(img)} bytes, {seg.info().n_row_groups} row groups", flush=True)
        staged += session.stage_segment(t, seg)
        seg.close()
    assert staged == len(t.row_groups) * len(t.schema)
    for req in (Request(["k1"], ALL_AGGS),
                Request(["k1", "k2"], ALL_AGGS, [SortSpec("sb", True), SortSpec("k1"), SortSpec("k2")], 25),
                Request(["k2"], ALL_AGGS, where_filter=Predicate("t", "lt", 4 * 10 ** 11)),
                Request(["k1"], [AggSpec("count", None, "c"), AggSpec("count_distinct", "s", "ds")],
                        having_filter=Predicate("c", "gt", 190))):
        check_vs_oracle(session, oracle_lib, t, req)
    # every block came from the files: staging the table again copies nothing
    assert session.stage_table(t, [c.name for c in t.schema]) == 0
