"""Segment files on the host (no device): the library's reader of thinDB's
.dat format (src/storage/segment_reader.zig:1350-1512, block payloads
:1577-1716, int_streams.zig) against files written by tests/segment_file.py.

Pins: every block's payload read back equals the payload written, across
raw / FOR / RLE bodies, nullable columns, no / zstd / LZ4 compression and
every int_streams form (plain, frame, delta, byte planes); int_streams
vectors built by hand from the format's stated formulas (frame:
base + stored * factor; delta: running sums from `first`; planes: byte b of
every value together) restore to the numbers written in the test; the
footer's stats and sizes read back; malformed files are rejected.  The
byte layout is pinned against the reference's reader and writer source;
no segment file written by thinDB itself exists to read (parity with its
writer's exact bytes is unpinned)."""
import struct

import numpy as np
import pytest

from segment_file import (BlockSpec, SegmentColumn, Stream, block_bytes, payload_of, stats_of, transform,
                          write_segment)
from thindb_amd import gpu
from thindb_amd.gpu import Type


def _cols(rng, n):
    v32 = rng.integers(-1000, 1000, n).astype(np.int32)
    runs = np.repeat(rng.integers(-5, 5, n // 50 + 1), 50)[:n].astype(np.int64)
    big = rng.integers(-2 ** 62, 2 ** 62, n, dtype=np.int64)
    seq = (np.arange(n, dtype=np.int64) * 7 + 10 ** 12)
    d = rng.uniform(-5, 5, n)
    ok = rng.random(n) > 0.2
    return [
        (Type.INT, v32, None, False),
        (Type.BIGINT, runs, ok, True),
        (Type.BIGINT, big, None, False),
        (Type.DATETIME, seq, ok, True),
        (Type.SMALLINT, (v32 // 10).astype(np.int16), None, False),
        (Type.DOUBLE, d, ok, True),
    ]


SPECS = [BlockSpec(enc, comp, form, shuf)
         for enc in ("raw", "for", "rle")
         for comp in ("none", "zstd", "lz4")
         for form, shuf in ((None, False), ("plain", False), ("frame", False), ("delta", False), ("frame", True),
                            ("delta", True), ("plain", True))
         if not (comp == "none" and form is not None and enc == "rle" and shuf)]


def test_segment_round_trip_every_form():
    """Every (encoding, compression, int_streams form) block reads back to
    the payload that was written, and the footer's stats read back."""
    rng = np.random.default_rng(0)
    n = 3000
    cols = _cols(rng, n)
    for spec in SPECS:
        segc = []
        for typ, v, ok, nullable in cols:
            s = spec if (typ not in (Type.FLOAT, Type.DOUBLE) or spec.encoding == "raw") else \
                BlockSpec("raw", spec.compression, spec.int_streams, spec.shuffled)
            segc.append(SegmentColumn(typ, v, ok, nullable, [s]))
        img = write_segment(segc, segment_id=42, rg_rows=1024)
        seg = gpu.Segment(img, len(segc))
        info = seg.info()
        assert (info.segment_id, info.row_count, info.n_row_groups, info.n_columns) == (42, n, 3, len(segc))
        assert info.schema_fingerprint == 0x5EED
        for rg, r0 in enumerate(range(0, n, 1024)):
            r1 = min(n, r0 + 1024)
            for ci, c in enumerate(segc):
                val = None if c.validity is None else c.validity[r0:r1]
                want, _ = payload_of(c.type, c.values[r0:r1], val, c.nullable, c.specs[0].encoding)
                b = seg.block(rg, ci)
                assert (b.row_count, b.payload_bytes, b.has_nulls) == (r1 - r0, len(want), int(c.nullable))
                assert b.compression == {"none": 0, "zstd": 1, "lz4": 2}[c.specs[0].compression]
                assert b.encoding == {"raw": 0, "for": 1, "rle": 3}[c.specs[0].encoding]
                assert seg.read_block(rg, ci) == want, (spec, rg, ci)
                mn, mx, sm, nulls = stats_of(c.type, c.values[r0:r1], val)
                assert (b.min, b.max, b.sum, b.null_count) == (mn, mx, sm, nulls)
        seg.close()


def _block_with(entries_and_stored: bytes, payload_len: int, compression: int = 0) -> bytes:
    return struct.pack("<BBBBII", compression, 2, 0, 0, payload_len, len(entries_and_stored)) + entries_and_stored


def _one_block_segment(block: bytes, rows: int) -> bytes:
    head = b"tDBS" + struct.pack("<HHQQQ", 17, 0, 0, 7, rows)
    rg_head = struct.pack("<II", rows, 0)
    body = head + rg_head
    off = len(body)
    body += block
    footer = struct.pack("<I", 1) + struct.pack("<QII", len(head), len(rg_head) + len(block), rows)
    footer += bytes(48) + struct.pack("<Q", 0) + struct.pack("<Q", off) + struct.pack("<I", len(block))
    footer += struct.pack("<I", 0)
    return body + footer + struct.pack("<I", len(footer) + 8) + b"tDBS"


def _entry(offset, count, width, method, stored_width, base=0, factor=1, first=0, shuffled=False):
    return struct.pack("<IIBBBBQQQ", offset, count, width, method | (0x80 if shuffled else 0), stored_width, 0,
                       base % 2 ** 64, factor % 2 ** 64, first % 2 ** 64)


def test_int_streams_known_answers():
    """Hand-built int_streams blocks, restored to the numbers the format's
    formulas give: frame 100 + s*3; delta from 5 by 2 with zero-width
    stored values; byte planes; the bytes after the last array kept."""
    # frame: four u16 values 100, 103, 106, 109 stored as one byte each, then a tail
    d = struct.pack("<I", 1) + _entry(0, 4, 2, 1, 1, base=100, factor=3)
    blk = _block_with(d + bytes([0, 1, 2, 3]) + b"TAIL", 8 + 4)
    seg = gpu.Segment(_one_block_segment(blk, 4), 1)
    assert seg.read_block(0, 0) == struct.pack("<4H", 100, 103, 106, 109) + b"TAIL"
    # delta: u32 values 5, 7, 9, 11 (first 5, base 2, stored width 0)
    d = struct.pack("<I", 1) + _entry(0, 4, 4, 2, 0, base=2, first=5)
    seg = gpu.Segment(_one_block_segment(_block_with(d, 16), 4), 1)
    assert seg.read_block(0, 0) == struct.pack("<4I", 5, 7, 9, 11)
    # delta with a negative step: i64 values 10, 7, 4 (base -3 two's complement)
    d = struct.pack("<I", 1) + _entry(0, 3, 8, 2, 0, base=-3, first=10)
    seg = gpu.Segment(_one_block_segment(_block_with(d, 24), 3), 1)
    assert seg.read_block(0, 0) == struct.pack("<3q", 10, 7, 4)
    # byte planes: plain u16 values 0x0102, 0x0304 stored as planes [02 04][01 03]
    d = struct.pack("<I", 1) + _entry(0, 2, 2, 0, 2, shuffled=True)
    seg = gpu.Segment(_one_block_segment(_block_with(d + bytes([2, 4, 1, 3]), 4), 2), 1)
    assert seg.read_block(0, 0) == struct.pack("<2H", 0x0102, 0x0304)
    # two arrays with a gap: [u8 tag][u16 frame x3][u8 tag] → tags kept, values 50, 60, 70
    d = struct.pack("<I", 1) + _entry(1, 3, 2, 1, 1, base=50, factor=10)
    seg = gpu.Segment(_one_block_segment(_block_with(d + b"A" + bytes([0, 1, 2]) + b"Z", 8), 3), 1)
    assert seg.read_block(0, 0) == b"A" + struct.pack("<3H", 50, 60, 70) + b"Z"


def test_transform_matches_hand_forms():
    """The test writer's frame/delta forms of a known array: their entries
    carry the parameters the formulas need (pins the writer the round trip
    uses)."""
    arr = np.array([1000, 1010, 1020, 1050], dtype=np.int32).tobytes()
    d, stored = transform(arr, [Stream(0, 4, 4, True)], "frame", False)
    n, = struct.unpack_from("<I", d)
    off, cnt, w, m, sw, _, base, factor, first = struct.unpack_from("<IIBBBBQQQ", d, 4)
    assert (n, off, cnt, w, m, sw, base, factor) == (1, 0, 4, 4, 1, 1, 1000, 10) and stored == bytes([0, 1, 2, 5])
    d, stored = transform(arr, [Stream(0, 4, 4, True)], "delta", False)
    _, _, _, m, sw, _, base, factor, first = struct.unpack_from("<IIBBBBQQQ", d, 4)
    # steps 10, 10, 30: base 10, then 0, 0, 20 over their common factor 20
    assert (m, sw, base, factor, first) == (2, 1, 10, 20, 1000) and stored == bytes([0, 0, 1])


@pytest.mark.parametrize("case", ["magic", "version", "footer_magic", "footer_size", "truncated", "block_range",
                                  "entry_width", "rows", "zstd_garbage"])
def test_segment_rejects_malformed(case):
    rng = np.random.default_rng(1)
    v = rng.integers(0, 100, 500).astype(np.int64)
    img = bytearray(write_segment([SegmentColumn(Type.BIGINT, v, specs=[BlockSpec("raw", "zstd", "frame")])], 3, 256))
    code = gpu.Status.INVALID
    if case == "magic":
        img[0:4] = b"XXXX"
    elif case == "version":
        img[4:6] = struct.pack("<H", 16)
        code = gpu.Status.UNSUPPORTED
    elif case == "footer_magic":
        img[-4:] = b"XXXX"
    elif case == "footer_size":
        img[-8:-4] = struct.pack("<I", len(img) + 5)
    elif case == "truncated":
        img = img[:20]
    elif case == "rows":
        img[24:32] = struct.pack("<Q", 499)
    else:
        footer_size, = struct.unpack("<I", img[-8:-4])
        f0 = len(img) - footer_size
        col_off_at = f0 + 4 + 16 + 56
        off, = struct.unpack_from("<Q", img, col_off_at)
        if case == "block_range":
            img[col_off_at:col_off_at + 8] = struct.pack("<Q", len(img) - 4)
        elif case == "entry_width":
            img[off + 12 + 4 + 10] = 9  # stored width 9
        else:
            img[off + 12 + 4 + 36: off + 12 + 4 + 44] = b"\xff" * 8  # the zstd frame's magic
    try:
        seg = gpu.Segment(bytes(img), 1)
    except gpu.GpuError as e:
        assert e.status == code, (case, e)
        return
    with pytest.raises(gpu.GpuError) as ei:
        seg.read_block(0, 0)
    assert ei.value.status == code, case


def test_block_bytes_header():
    """A block's 12-byte header: compression, flags (has_nulls, int_streams),
    encoding, reserved, payload size, stored size."""
    v = np.arange(300, dtype=np.int32)
    b = block_bytes(Type.INT, v, np.ones(300, bool), True, BlockSpec("for", "lz4", "delta", True))
    comp, flags, enc, res, usize, csize = struct.unpack_from("<BBBBII", b)
    assert (comp, flags, enc, res) == (2, 3, 1, 0) and csize == len(b) - 12
    assert usize == len(payload_of(Type.INT, v, np.ones(300, bool), True, "for")[0])


def test_segment_survives_corruption():
    """Random byte flips and truncations of a valid file: opening and
    reading either succeed or fail with an error status; every offset the
    reader follows is checked first (a crash would end the test process)."""
    rng = np.random.default_rng(5)
    v = rng.integers(-50, 50, 2000).astype(np.int64)
    base = write_segment([SegmentColumn(Type.BIGINT, v, rng.random(2000) > 0.1, True,
                                        [BlockSpec("for", "zstd", "delta", True), BlockSpec("rle", "lz4", "frame"),
                                         BlockSpec("raw", "none", "plain")]),
                          SegmentColumn(Type.INT, v.astype(np.int32), specs=[BlockSpec("raw", "zstd", "frame")])],
                         9, 512)
    outcomes = {"ok": 0, "error": 0}
    for trial in range(400):
        img = bytearray(base)
        if trial % 5 == 4:
            img = img[:int(rng.integers(0, len(img)))]
        else:
            for _ in range(int(rng.integers(1, 8))):
                img[int(rng.integers(0, len(img)))] = int(rng.integers(0, 256))
        try:
            seg = gpu.Segment(bytes(img), 2)
            info = seg.info()
            for rg in range(info.n_row_groups):
                for ci in range(2):
                    seg.read_block(rg, ci)
            seg.close()
            outcomes["ok"] += 1
        except gpu.GpuError as e:
            assert e.status in (gpu.Status.INVALID, gpu.Status.UNSUPPORTED), e
            outcomes["error"] += 1
    assert outcomes["error"] > 0 and outcomes["ok"] > 0


def test_segment_parser_under_asan(tmp_path):
    """The parser built for the host with AddressSanitizer and UBSan
    (tests/native/segment_fuzz.cpp; the device staging call stubbed) reads
    4,000 corrupted and truncated copies of a segment file without a
    sanitizer report."""
    import shutil
    import subprocess
    from pathlib import Path
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    root = Path(__file__).resolve().parents[1]
    rng = np.random.default_rng(5)
    v = rng.integers(-50, 50, 2000).astype(np.int64)
    img = write_segment([SegmentColumn(Type.BIGINT, v, rng.random(2000) > 0.1, True,
                                       [BlockSpec("for", "zstd", "delta", True), BlockSpec("rle", "lz4", "frame"),
                                        BlockSpec("raw", "none", "plain")]),
                         SegmentColumn(Type.INT, v.astype(np.int32), specs=[BlockSpec("raw", "zstd", "frame")])],
                        9, 512)
    (tmp_path / "seg.dat").write_bytes(img)
    exe = tmp_path / "fuzz"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{root / 'thindb_amd' / 'csrc'}", str(root / "thindb_amd" / "csrc" / "segment.cpp"),
                    str(root / "tests" / "native" / "segment_fuzz.cpp"), "-o", str(exe), "-ldl"], check=True)
    out = subprocess.run([str(exe), str(tmp_path / "seg.dat"), "4000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert "ERROR: AddressSanitizer" not in out.stderr and "runtime error" not in out.stderr
