"""Writes thinDB segment files (.dat, format version 17) for the tests.

The layout follows the format as the reference's reader parses it and its
writer lays it out: the 32-byte header (headerBytes,
src/storage/segment_writer.zig:90-99), per row group an 8-byte header and one
column block per column (appendRowGroupBlocks :359-390), each block a 12-byte
header and its stored bytes (appendBlock :1196-1250), then the footer
(appendFooter :642-680) and the 8-byte trailer.  A block's payload is the
validity bitmap when the column is nullable, then its raw / FOR / RLE body
(writeColumnBlock :1141-1190).  A compressed block may rewrite its integer
arrays first (int_streams: plain, frame or delta form, optionally split into
byte planes; src/storage/int_streams.zig:1-40); the form is chosen by the
caller here, where the writer picks it by a zstd trial.  Compression uses the
system's libzstd.so.1 / liblz4.so.1 and is applied whether or not it
shrinks the block (the writer keeps it only when it does; the reader reads
both).
"""
import ctypes as C
import math
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from helpers import bitmap_bytes, for_payload, rle_payload
from thindb_amd.gpu import NUMPY_DTYPE, Type

MAGIC = b"tDBS"
VERSION = 17
COMPRESSION = {"none": 0, "zstd": 1, "lz4": 2}
ENCODING = {"raw": 0, "for": 1, "rle": 3}


def _zstd():
    lib = C.CDLL("libzstd.so.1")
    lib.ZSTD_compressBound.restype = C.c_size_t
    lib.ZSTD_compressBound.argtypes = [C.c_size_t]
    lib.ZSTD_compress.restype = C.c_size_t
    lib.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
    lib.ZSTD_isError.restype = C.c_uint
    lib.ZSTD_isError.argtypes = [C.c_size_t]
    return lib


def _lz4():
    lib = C.CDLL("liblz4.so.1")
    lib.LZ4_compressBound.restype = C.c_int
    lib.LZ4_compressBound.argtypes = [C.c_int]
    lib.LZ4_compress_default.restype = C.c_int
    lib.LZ4_compress_default.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int]
    return lib


def compress(kind: str, data: bytes) -> bytes:
    if kind == "none":
        return data
    if kind == "zstd":
        z = _zstd()
        cap = z.ZSTD_compressBound(len(data))
        out = C.create_string_buffer(cap)
        n = z.ZSTD_compress(out, cap, data, len(data), 6)
        assert not z.ZSTD_isError(n)
        return out.raw[:n]
    lz = _lz4()
    cap = lz.LZ4_compressBound(len(data))
    out = C.create_string_buffer(cap)
    n = lz.LZ4_compress_default(data, out, len(data), cap)
    assert n > 0
    return out.raw[:n]


# ---- int_streams forms -------------------------------------------------------------------

@dataclass
class Stream:
    offset: int
    count: int
    width: int
    signed: bool


def _load(payload: bytes, s: Stream) -> List[int]:
    out = []
    for i in range(s.count):
        v = int.from_bytes(payload[s.offset + i * s.width: s.offset + (i + 1) * s.width], "little", signed=s.signed)
        out.append(v & (2 ** 64 - 1))
    return out


def _narrow(mx: int) -> int:
    for w in (0, 1, 2, 4, 8):
        if mx < 1 << (8 * w):
            return w
    return 8


def _signed64(v: int) -> int:
    return v - 2 ** 64 if v >> 63 else v


def rewrite(payload: bytes, s: Stream, method: str, shuffled: bool):
    """(entry bytes, stored bytes) of one array in the given form."""
    M = 2 ** 64
    vals = _load(payload, s)
    base = factor = first = 0
    factor = 1
    if method == "plain":
        stored, sw = vals, s.width
        mcode = 0
    elif method == "frame":
        keyed = [_signed64(v) if s.signed else v for v in vals]
        base = min(keyed) % M
        diffs = [(v - base) % M for v in vals]
        factor = math.gcd(*diffs) or 1
        stored = [d // factor for d in diffs]
        sw = min(_narrow(max(stored)), s.width)
        mcode = 1
    else:
        first = vals[0]
        d = [(vals[i] - vals[i - 1]) % M for i in range(1, len(vals))]
        base = (min(_signed64(x) for x in d) % M) if d else 0
        e = [(x - base) % M for x in d]
        factor = (math.gcd(*e) if e else 0) or 1
        stored = [x // factor for x in e]
        sw = min(_narrow(max(stored)) if stored else 0, s.width)
        mcode = 2
    if sw:
        if shuffled:
            body = b"".join(bytes((v >> (8 * b)) & 0xFF for v in stored) for b in range(sw))
        else:
            body = b"".join((v & ((1 << (8 * sw)) - 1)).to_bytes(sw, "little") for v in stored)
    else:
        body = b""
    entry = struct.pack("<IIBBBBQQQ", s.offset, s.count, s.width, mcode | (0x80 if shuffled else 0), sw, 0,
                        base % M, factor % M, first % M)
    return entry, body


def transform(payload: bytes, streams: Sequence[Stream], method: str, shuffled: bool):
    """The directory and stored payload with every array rewritten."""
    entries, out, at = [], b"", 0
    for s in streams:
        e, body = rewrite(payload, s, method, shuffled)
        entries.append(e)
        out += payload[at:s.offset] + body
        at = s.offset + s.count * s.width
    out += payload[at:]
    return struct.pack("<I", len(entries)) + b"".join(entries), out


# ---- blocks ------------------------------------------------------------------------------

@dataclass
class BlockSpec:
    encoding: str = "raw"          # raw | for | rle
    compression: str = "none"      # none | zstd | lz4
    int_streams: Optional[str] = None  # None | plain | frame | delta
    shuffled: bool = False


def payload_of(typ: Type, values: np.ndarray, validity, nullable: bool, encoding: str):
    """(payload bytes, its integer arrays) of one block."""
    dt = np.dtype(NUMPY_DTYPE[typ])
    v = np.ascontiguousarray(values, dtype=dt)
    ok = np.ones(len(v), bool) if validity is None else np.asarray(validity, bool)
    bm = bitmap_bytes(ok) if nullable else b""
    signed = typ not in (Type.FLOAT, Type.DOUBLE)
    if encoding == "raw":
        streams = [Stream(len(bm), len(v), dt.itemsize, True)] if signed else []
        return bm + v.tobytes(), streams
    if encoding == "for":
        p = for_payload(v, ok if nullable else None)
        width = p[len(bm) + 16]
        return p, [Stream(len(bm) + 20, len(v), width, False)]
    p = rle_payload(v, ok if nullable else None)
    n_runs = int.from_bytes(p[len(bm):len(bm) + 4], "little")
    w = dt.itemsize
    return p, [Stream(len(bm) + 8, n_runs, w, True), Stream(len(bm) + 8 + n_runs * w, n_runs, 4, False)]


def block_bytes(typ, values, validity, nullable, spec: BlockSpec) -> bytes:
    payload, streams = payload_of(typ, values, validity, nullable, spec.encoding)
    flags = 1 if nullable else 0
    directory = b""
    source = payload
    if spec.int_streams and streams:
        directory, source = transform(payload, streams, spec.int_streams, spec.shuffled)
        flags |= 2
    stored = directory + compress(spec.compression, source)
    head = struct.pack("<BBBBII", COMPRESSION[spec.compression], flags, ENCODING[spec.encoding], 0, len(payload),
                       len(stored))
    return head + stored


def float_order_key(f: float) -> int:
    """encodeFloatOrder (src/storage/format.zig:301-309)."""
    if f != f:
        return 2 ** 64 - 1
    bits = struct.unpack("<Q", struct.pack("<d", 0.0 if f == 0 else f))[0]
    return (~bits) & (2 ** 64 - 1) if bits >> 63 else bits | (1 << 63)


def stats_of(typ, values, validity):
    v = np.asarray(values)
    ok = np.ones(len(v), bool) if validity is None else np.asarray(validity, bool)
    nulls = int((~ok).sum())
    live = v[ok]
    if typ in (Type.FLOAT, Type.DOUBLE):
        f = [float(x) for x in live if x == x]
        mn = float_order_key(min(f)) if f else 0
        mx = float_order_key(max(f)) if f else 0
        sm = struct.unpack("<q", struct.pack("<d", float(np.sum(live.astype(np.float64)))))[0] if len(live) else 0
        return mn, mx, sm, nulls
    if len(live) == 0:
        return 0, 0, 0, nulls
    return int(live.min()), int(live.max()), int(sum(int(x) for x in live)), nulls


@dataclass
class SegmentColumn:
    type: Type
    values: np.ndarray
    validity: Optional[np.ndarray] = None
    nullable: bool = False
    specs: List[BlockSpec] = field(default_factory=list)  # per row group (cycled); default raw/none


def write_segment(cols: Sequence[SegmentColumn], segment_id: int, rg_rows: int, fingerprint: int = 0x5EED) -> bytes:
    n = len(cols[0].values)
    out = bytearray(MAGIC + struct.pack("<HHQQQ", VERSION, 0, fingerprint, segment_id, n))
    rgs = []
    for gi, r0 in enumerate(range(0, n, rg_rows)):
        r1 = min(n, r0 + rg_rows)
        off = len(out)
        out += struct.pack("<II", r1 - r0, 0)
        col_off, col_len, stats = [], [], []
        for c in cols:
            spec = c.specs[gi % len(c.specs)] if c.specs else BlockSpec()
            val = None if c.validity is None else c.validity[r0:r1]
            b = block_bytes(c.type, c.values[r0:r1], val, c.nullable, spec)
            col_off.append(len(out))
            col_len.append(len(b))
            out += b
            stats.append(stats_of(c.type, c.values[r0:r1], val))
        rgs.append((off, len(out) - off, r1 - r0, stats, col_off, col_len))
    footer = bytearray(struct.pack("<I", len(rgs)))
    for off, length, rows, stats, col_off, col_len in rgs:
        footer += struct.pack("<QII", off, length, rows)
        for mn, mx, sm, nulls in stats:
            footer += mn.to_bytes(16, "little", signed=True) + mx.to_bytes(16, "little", signed=True)
            footer += sm.to_bytes(16, "little", signed=True) + struct.pack("<Q", nulls)
        footer += b"".join(struct.pack("<Q", o) for o in col_off)
        footer += b"".join(struct.pack("<I", l) for l in col_len)
    footer += struct.pack("<I", 0)  # no string tables
    out += footer + struct.pack("<I", len(footer) + 8) + MAGIC
    return bytes(out)
