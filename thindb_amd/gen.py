"""Seeded synthetic tables for the GROUP BY workloads (binding of libthindb_gen.so).

Every value is a pure function of (seed, column, row), so any slice of a
table can be regenerated independently and identically.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Dict, Tuple

import numpy as np

from .exec import Column, Table
from .gpu import Type

_LIB = Path(__file__).resolve().parent / "libthindb_gen.so"
_lib = None

T0_US = 1767225600000000          # 2026-01-01 00:00:00 UTC in microseconds
DAY_US = 86400 * 1000000
FILTER_LO = T0_US + 10 * DAY_US    # WHERE ts >= T0 + 10 days
FILTER_HI = T0_US + 17 * DAY_US    #   AND ts <  T0 + 17 days (7/30 of rows)


def _threads() -> int:
    n = os.cpu_count() or 8
    return max(1, min(16, n))


def lib():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            raise RuntimeError(f"{_LIB} is not built: run __graft_entry__.build()")
        L = C.CDLL(str(_LIB))
        p64 = C.POINTER(C.c_int64)
        L.tdbg_gen_plain.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, p64, p64, C.c_int]
        L.tdbg_gen_filtered.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_int32), p64, p64,
                                        C.POINTER(C.c_double), C.c_int]
        L.tdbg_gen_highcard.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, p64, C.POINTER(C.c_int8),
                                        C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int]
        L.tdbg_gen_batched_shape.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64), p64,
                                             C.POINTER(C.c_int32)]
        L.tdbg_gen_batched.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(C.c_int32), p64,
                                       C.c_int]
        L.tdbg_gen_keyed.argtypes = [C.c_uint64, C.c_int64, C.c_uint64, C.c_uint64, C.POINTER(C.c_int32), p64,
                                     C.c_int]
        L.tdbg_gen_zipf_pmf.argtypes = [C.c_uint32, C.c_double, C.c_uint32]
        L.tdbg_gen_zipf_pmf.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def plain(n: int, seed: int = 1, row0: int = 0) -> Dict[str, np.ndarray]:
    """k BIGINT U[0,1024), v BIGINT U[-1e6, 1e6]."""
    k = np.empty(n, np.int64)
    v = np.empty(n, np.int64)
    lib().tdbg_gen_plain(seed, row0, n, _ptr(k, C.c_int64), _ptr(v, C.c_int64), _threads())
    return {"k": k, "v": v}


PLAIN_SCHEMA = [Column("k", Type.BIGINT), Column("v", Type.BIGINT)]


def filtered(n: int, seed: int = 2, row0: int = 0) -> Dict[str, np.ndarray]:
    """region INT Zipf(1.1, 1e4); ts DATETIME over 30 days; amount DECIMAL(18,2)
    log-normal(8, 2) cents clipped to 1e12; price DOUBLE U[0, 1000)."""
    region = np.empty(n, np.int32)
    ts = np.empty(n, np.int64)
    amount = np.empty(n, np.int64)
    price = np.empty(n, np.float64)
    lib().tdbg_gen_filtered(seed, row0, n, _ptr(region, C.c_int32), _ptr(ts, C.c_int64), _ptr(amount, C.c_int64),
                            _ptr(price, C.c_double), _threads())
    return {"region": region, "ts": ts, "amount": amount, "price": price}


FILTERED_SCHEMA = [Column("region", Type.INT), Column("ts", Type.DATETIME), Column("amount", Type.DECIMAL64, scale=2),
                   Column("price", Type.DOUBLE)]


def highcard(n: int, seed: int = 3, row0: int = 0) -> Dict[str, np.ndarray]:
    """user_id BIGINT U[0, 2e7); bucket TINYINT U[0,16); v INT U[0,1e5); w DOUBLE N(0,1)."""
    uid = np.empty(n, np.int64)
    bucket = np.empty(n, np.int8)
    v = np.empty(n, np.int32)
    w = np.empty(n, np.float64)
    lib().tdbg_gen_highcard(seed, row0, n, _ptr(uid, C.c_int64), _ptr(bucket, C.c_int8), _ptr(v, C.c_int32),
                            _ptr(w, C.c_double), _threads())
    return {"user_id": uid, "bucket": bucket, "v": v, "w": w}


HIGHCARD_SCHEMA = [Column("user_id", Type.BIGINT), Column("bucket", Type.TINYINT), Column("v", Type.INT),
                   Column("w", Type.DOUBLE)]


def wide(n: int, seed: int = 4) -> Dict[str, np.ndarray]:
    """Two full-range BIGINT keys (a, b): n // 2 random pairs over all 2^64
    values of each, every row picking one uniformly (about 43% of n distinct
    pairs); v INT uniform in [-1000, 1000).  The shape of GROUP BY WatchID,
    ClientIP: no relative packing narrows the key below 128 bits."""
    rng = np.random.default_rng(seed)
    lo, hi = np.iinfo(np.int64).min, np.iinfo(np.int64).max
    pairs = max(1, n // 2)
    ga = rng.integers(lo, hi, pairs, dtype=np.int64, endpoint=True)
    gb = rng.integers(lo, hi, pairs, dtype=np.int64, endpoint=True)
    pick = rng.integers(0, pairs, n)
    return {"a": ga[pick], "b": gb[pick], "v": rng.integers(-1000, 1000, n).astype(np.int32)}


WIDE_SCHEMA = [Column("a", Type.BIGINT), Column("b", Type.BIGINT), Column("v", Type.INT)]


def batched_shape(seed: int, t: int) -> Tuple[int, int, bool]:
    """(rows, key cardinality, filtered) of batched table t."""
    rows = C.c_uint64()
    card = C.c_int64()
    f = C.c_int32()
    lib().tdbg_gen_batched_shape(seed, t, C.byref(rows), C.byref(card), C.byref(f))
    return int(rows.value), int(card.value), bool(f.value)


def batched(seed: int, t: int, n: int, row0: int = 0) -> Dict[str, np.ndarray]:
    """key INT U[0, card); value BIGINT U[-1e9, 1e9]."""
    key = np.empty(n, np.int32)
    value = np.empty(n, np.int64)
    lib().tdbg_gen_batched(seed, t, row0, n, _ptr(key, C.c_int32), _ptr(value, C.c_int64), _threads())
    return {"key": key, "value": value}


BATCHED_SCHEMA = [Column("key", Type.INT), Column("value", Type.BIGINT)]


def zipf_pmf(n: int, s: float, k: int) -> float:
    return float(lib().tdbg_gen_zipf_pmf(n, s, k))


def table(kind: str, n: int, seed: int = None, **kw) -> Table:
    """A host Table of one workload kind: plain, filtered, highcard or wide."""
    if kind == "plain":
        return Table(PLAIN_SCHEMA, plain(n, seed if seed is not None else 1), **kw)
    if kind == "filtered":
        return Table(FILTERED_SCHEMA, filtered(n, seed if seed is not None else 2), **kw)
    if kind == "highcard":
        return Table(HIGHCARD_SCHEMA, highcard(n, seed if seed is not None else 3), **kw)
    if kind == "wide":
        return Table(WIDE_SCHEMA, wide(n, seed if seed is not None else 4), **kw)
    raise ValueError(kind)


BATCH_FILTER_LO = -500_000_000   # optional filter: value >= -5e8 AND value < 5e8 (50% of rows)
BATCH_FILTER_HI = 500_000_000


def batched_table(seed: int, t: int, max_rows: int = 0) -> Table:
    """Batched-workload table t (rows log-uniform in [1e4, 2e7], key
    cardinality log-uniform in [4, 1e6]); max_rows > 0 truncates it."""
    rows, _, _ = batched_shape(seed, t)
    n = min(rows, max_rows) if max_rows else rows
    return Table(BATCHED_SCHEMA, batched(seed, t, n))


def batched_request(seed: int, t: int):
    """GROUP BY key with COUNT/SUM/MIN/MAX of value, with the 50%-selective
    range filter on value for the tables whose coin flip says so."""
    from .exec import AggSpec, And, Predicate, Request
    _, _, filtered_ = batched_shape(seed, t)
    where = And((Predicate("value", "gte", BATCH_FILTER_LO), Predicate("value", "lt", BATCH_FILTER_HI))) \
        if filtered_ else None
    return Request(["key"], [AggSpec("count", None, "c"), AggSpec("sum", "value", "s"), AggSpec("min", "value", "mn"),
                             AggSpec("max", "value", "mx")], where_filter=where)


def keyed_table(seed: int, n: int, card: int) -> Table:
    """n rows of the batched-workload shape with key cardinality `card`."""
    key = np.empty(n, np.int32)
    value = np.empty(n, np.int64)
    lib().tdbg_gen_keyed(seed, card, 0, n, _ptr(key, C.c_int32), _ptr(value, C.c_int64), _threads())
    return Table(BATCHED_SCHEMA, {"key": key, "value": value})


# ---- segment files -------------------------------------------------------------------------
# A table written as thinDB segment files (format version 17), for measuring
# ingestion: the header, row groups of column blocks and the footer as the
# reference's writer lays them out (src/storage/segment_writer.zig:90-99,
# :359-390, :642-680, block header :1196-1250).  Integer columns take the
# frame-of-reference body when it is narrower than the native width (base =
# min, codes = value - min; tryEncodeFor, :1254-1290), else raw; floats are
# raw; a block keeps its zstd frame only when that is smaller (appendBlock).

def _zstd():
    z = C.CDLL("libzstd.so.1")
    z.ZSTD_compressBound.restype = C.c_size_t
    z.ZSTD_compressBound.argtypes = [C.c_size_t]
    z.ZSTD_compress.restype = C.c_size_t
    z.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
    z.ZSTD_isError.restype = C.c_uint
    z.ZSTD_isError.argtypes = [C.c_size_t]
    return z


def _float_order_key(f: float) -> int:
    """encodeFloatOrder (src/storage/format.zig:301-309)."""
    bits = int(np.array([0.0 if f == 0 else f], dtype=np.float64).view(np.uint64)[0])
    return (~bits) & (2 ** 64 - 1) if bits >> 63 else bits | (1 << 63)


def _segment_block(col: Column, v: np.ndarray, zstd, level: int):
    """(block bytes, (min, max, sum, null_count)) of one non-nullable block."""
    if col.type in (Type.FLOAT, Type.DOUBLE):
        enc, payload = 0, np.ascontiguousarray(v).tobytes()
        live = v[~np.isnan(v)]
        mn = _float_order_key(float(live.min())) if len(live) else 0
        mx = _float_order_key(float(live.max())) if len(live) else 0
        sm = int(np.array([v.astype(np.float64).sum()]).view(np.int64)[0])
    else:
        lo, hi = int(v.min()), int(v.max())
        span = hi - lo
        width = 1 if span < 1 << 8 else 2 if span < 1 << 16 else 4 if span < 1 << 32 else 8
        mn, mx, sm = lo, hi, int(v.astype(np.int64).sum(dtype=np.int64))
        if width < v.dtype.itemsize:
            codes = (v.astype(np.int64) - lo).astype(np.dtype(f"<u{width}"))
            enc = 1
            payload = lo.to_bytes(16, "little", signed=True) + bytes([width, 0, 0, 0]) + codes.tobytes()
        else:
            enc, payload = 0, np.ascontiguousarray(v).tobytes()
    comp, stored = 0, payload
    if zstd is not None:
        cap = zstd.ZSTD_compressBound(len(payload))
        out = C.create_string_buffer(cap)
        n = zstd.ZSTD_compress(out, cap, payload, len(payload), level)
        if not zstd.ZSTD_isError(n) and n < len(payload):
            comp, stored = 1, out.raw[:n]
    head = int(comp).to_bytes(1, "little") + b"\x00" + bytes([enc, 0]) + len(payload).to_bytes(4, "little") + \
        len(stored).to_bytes(4, "little")
    return head + stored, (mn, mx, sm, 0)


def segment_images(t: Table, compression: str = "zstd", level: int = 3, threads: int = 0) -> Dict[int, bytes]:
    """Every segment of `t` (no NULLs) as a file image, by segment id."""
    from concurrent.futures import ThreadPoolExecutor

    zstd = _zstd() if compression == "zstd" else None
    by_seg: Dict[int, list] = {}
    for rg in t.row_groups:
        by_seg.setdefault(rg.segment_id, []).append(rg)

    def row_group(rg):
        return [_segment_block(c, t.columns[c.name][rg.start:rg.start + rg.n_rows], zstd, level) for c in t.schema]

    out = {}
    with ThreadPoolExecutor(threads or _threads()) as ex:
        for seg_id, rgs in by_seg.items():
            blocks = list(ex.map(row_group, rgs))
            n = sum(rg.n_rows for rg in rgs)
            img = bytearray(b"tDBS" + (17).to_bytes(2, "little") + bytes(2) + bytes(8) +
                            int(seg_id).to_bytes(8, "little") + n.to_bytes(8, "little"))
            footer = bytearray(len(rgs).to_bytes(4, "little"))
            for rg, cols in zip(rgs, blocks):
                off = len(img)
                img += rg.n_rows.to_bytes(4, "little") + bytes(4)
                offs = []
                for b, _ in cols:
                    offs.append((len(img), len(b)))
                    img += b
                footer += off.to_bytes(8, "little") + (len(img) - off).to_bytes(4, "little") + \
                    rg.n_rows.to_bytes(4, "little")
                for _, (mn, mx, sm, nulls) in cols:
                    footer += b"".join(x.to_bytes(16, "little", signed=True) for x in (mn, mx, sm))
                    footer += nulls.to_bytes(8, "little")
                footer += b"".join(o.to_bytes(8, "little") for o, _ in offs)
                footer += b"".join(l.to_bytes(4, "little") for _, l in offs)
            footer += bytes(4)  # no string tables
            img += footer + (len(footer) + 8).to_bytes(4, "little") + b"tDBS"
            out[seg_id] = bytes(img)
    return out
