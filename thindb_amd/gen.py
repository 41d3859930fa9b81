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
    """A host Table of one workload kind: plain, filtered or highcard."""
    if kind == "plain":
        return Table(PLAIN_SCHEMA, plain(n, seed if seed is not None else 1), **kw)
    if kind == "filtered":
        return Table(FILTERED_SCHEMA, filtered(n, seed if seed is not None else 2), **kw)
    if kind == "highcard":
        return Table(HIGHCARD_SCHEMA, highcard(n, seed if seed is not None else 3), **kw)
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
