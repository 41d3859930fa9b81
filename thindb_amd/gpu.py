"""ctypes binding of ``libthindb_gpu.so`` (the C-ABI in ``include/thindb_gpu.h``).

This is the FFI layer a host binds; :mod:`thindb_amd.exec` builds thinDB's
operator interface on top of it.  Loading fails loudly when the library is
missing: there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

_LIB_PATH = Path(os.environ.get("THINDB_GPU_LIB", Path(__file__).resolve().parent / "libthindb_gpu.so"))

ABI_VERSION = 3


class Status(enum.IntEnum):
    OK = 0
    PENDING = 1
    INVALID = -1
    UNSUPPORTED = -2
    OOM = -3
    CANCELLED = -4
    DEVICE = -5
    NOT_STAGED = -6


class Encoding(enum.IntEnum):
    """Block value encodings (src/storage/format.zig:166-172)."""
    RAW = 0
    FOR = 1
    DICT = 2
    RLE = 3
    FSST = 4


class Type(enum.IntEnum):
    """Physical column types (src/types.zig subset)."""
    BOOLEAN = 1
    TINYINT = 2
    SMALLINT = 3
    INT = 4
    BIGINT = 5
    DATE = 6
    DATETIME = 7
    DECIMAL64 = 8
    FLOAT = 9
    DOUBLE = 10
    DECIMAL128 = 11


NUMPY_DTYPE = {
    Type.BOOLEAN: np.uint8,
    Type.TINYINT: np.int8,
    Type.SMALLINT: np.int16,
    Type.INT: np.int32,
    Type.BIGINT: np.int64,
    Type.DATE: np.int32,
    Type.DATETIME: np.int64,
    Type.DECIMAL64: np.int64,
    Type.FLOAT: np.float32,
    Type.DOUBLE: np.float64,
}


class AggOp(enum.IntEnum):
    COUNT_STAR = 0
    COUNT = 1
    SUM = 2
    AVG = 3
    MIN = 4
    MAX = 5
    COUNT_DISTINCT = 6


class PredOp(enum.IntEnum):
    """src/exec/predicate.zig:33 PredicateOp."""
    EQ = 0
    NEQ = 1
    LT = 2
    LTE = 3
    GT = 4
    GTE = 5


MAX_HAVING = 32  # TDBG_MAX_HAVING
MAX_WHERE = 12   # TDBG_MAX_WHERE


class HavingKind(enum.IntEnum):
    """enum tdbg_having_kind: postfix HAVING nodes."""
    CMP = 0
    IS_NULL = 1
    IS_NOT_NULL = 2
    NOT = 3
    AND = 4
    OR = 5


# ---- errors (status -> the reference's error set) ---------------------------

class GpuError(RuntimeError):
    status = Status.DEVICE


class Unsupported(GpuError):
    """The shape is not covered: the caller declines to its CPU operator."""
    status = Status.UNSUPPORTED


class MemoryBudgetExceeded(GpuError):
    """src/memory.zig:31-35 — device memory or admission budget exhausted."""
    status = Status.OOM


class QueryCancelled(GpuError):
    status = Status.CANCELLED


class NotStaged(GpuError):
    status = Status.NOT_STAGED


class InvalidArgument(GpuError):
    status = Status.INVALID


_ERRORS = {
    Status.UNSUPPORTED: Unsupported,
    Status.OOM: MemoryBudgetExceeded,
    Status.CANCELLED: QueryCancelled,
    Status.NOT_STAGED: NotStaged,
    Status.INVALID: InvalidArgument,
}


# ---- C structs ------------------------------------------------------------------

class ColumnKey(C.Structure):
    _fields_ = [("table_uid", C.c_uint64), ("segment_id", C.c_uint64),
                ("row_group_idx", C.c_uint32), ("column_idx", C.c_uint32)]


class SegmentInfo(C.Structure):
    _fields_ = [("segment_id", C.c_uint64), ("row_count", C.c_uint64), ("schema_fingerprint", C.c_uint64),
                ("n_row_groups", C.c_uint32), ("n_columns", C.c_uint32)]


class SegmentBlock(C.Structure):
    _fields_ = [("row_count", C.c_uint32), ("payload_bytes", C.c_uint32), ("stored_bytes", C.c_uint32),
                ("compression", C.c_uint8), ("has_nulls", C.c_uint8), ("int_streams", C.c_uint8),
                ("encoding", C.c_uint8), ("null_count", C.c_uint64), ("min_lo", C.c_uint64), ("min_hi", C.c_uint64),
                ("max_lo", C.c_uint64), ("max_hi", C.c_uint64), ("sum_lo", C.c_uint64), ("sum_hi", C.c_uint64)]

    @staticmethod
    def _i128(lo: int, hi: int) -> int:
        v = (hi << 64) | lo
        return v - (1 << 128) if v >> 127 else v

    @property
    def min(self) -> int:
        return self._i128(self.min_lo, self.min_hi)

    @property
    def max(self) -> int:
        return self._i128(self.max_lo, self.max_hi)

    @property
    def sum(self) -> int:
        return self._i128(self.sum_lo, self.sum_hi)


class BlockStats(C.Structure):
    _fields_ = [("min_i", C.c_int64), ("max_i", C.c_int64), ("min_f", C.c_double), ("max_f", C.c_double),
                ("has_minmax", C.c_uint32), ("pad", C.c_uint32)]


class RgRef(C.Structure):
    _fields_ = [("segment_id", C.c_uint64), ("row_group_idx", C.c_uint32), ("pad", C.c_uint32)]


class KeyPart(C.Structure):
    _fields_ = [("column_idx", C.c_uint32), ("type", C.c_uint32)]


class AggSpecC(C.Structure):
    _fields_ = [("op", C.c_uint32), ("column_idx", C.c_int32), ("input_type", C.c_uint32),
                ("input_scale", C.c_uint8), ("nullable", C.c_uint8), ("pad", C.c_uint8 * 2)]


class PredicateC(C.Structure):
    _fields_ = [("column_idx", C.c_uint32), ("type", C.c_uint32), ("op", C.c_uint32), ("nullable", C.c_uint32),
                ("i64", C.c_int64), ("f64", C.c_double)]


class OrderSpecC(C.Structure):
    _fields_ = [("target", C.c_uint32), ("desc", C.c_uint32)]


class HavingNodeC(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("target", C.c_uint32), ("op", C.c_uint32), ("arity", C.c_uint32),
                ("lit_float", C.c_uint32), ("pad", C.c_uint32), ("lit_lo", C.c_int64), ("lit_hi", C.c_int64),
                ("lit_f64", C.c_double)]


class FilterNodeC(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("column_idx", C.c_uint32), ("type", C.c_uint32), ("op", C.c_uint32),
                ("arity", C.c_uint32), ("pad", C.c_uint32), ("i64", C.c_int64), ("f64", C.c_double)]


class PlanDesc(C.Structure):
    _fields_ = [("table_uid", C.c_uint64), ("n_keys", C.c_uint32), ("n_aggs", C.c_uint32),
                ("keys", C.POINTER(KeyPart)), ("aggs", C.POINTER(AggSpecC)),
                ("n_preds", C.c_uint32), ("n_order", C.c_uint32),
                ("preds", C.POINTER(PredicateC)), ("order", C.POINTER(OrderSpecC)),
                ("limit", C.c_uint64), ("offset", C.c_uint64), ("est_groups", C.c_uint64),
                ("n_having", C.c_uint32), ("n_where", C.c_uint32), ("having", C.POINTER(HavingNodeC)),
                ("where", C.POINTER(FilterNodeC))]


class ColumnView(C.Structure):
    _fields_ = [("type", C.c_uint32), ("scale", C.c_uint32), ("nullable", C.c_uint32), ("pad", C.c_uint32),
                ("data", C.c_void_p), ("validity", C.c_void_p), ("n_rows", C.c_uint64)]


class ExecStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("scan_ms", C.c_double), ("rows_scanned", C.c_uint64),
                ("bytes_scanned", C.c_uint64), ("n_groups", C.c_uint64), ("lds_mode", C.c_uint32),
                ("global_mode", C.c_uint32), ("specialized", C.c_uint32), ("lds_parts", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("scan_ms", C.c_double), ("rows", C.c_uint64), ("bytes", C.c_uint64),
                ("result_bytes", C.c_uint64), ("n_plans", C.c_uint32), ("n_items", C.c_uint32),
                ("graph", C.c_uint32), ("waves", C.c_uint32)]


class BatchRoute(C.Structure):
    _fields_ = [("strategy", C.c_uint32), ("lds_parts", C.c_uint32), ("radix_bits", C.c_uint32),
                ("items", C.c_uint32)]


# Every entry point include/thindb_gpu.h declares, with its ctypes signature.
EXPORTS = {
    "tdbg_abi_version": (C.c_uint32, []),
    "tdbg_last_error": (C.c_char_p, []),
    "tdbg_context_create": (C.c_int32, [C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "tdbg_context_destroy": (None, [C.c_void_p]),
    "tdbg_context_bytes_in_use": (C.c_uint64, [C.c_void_p]),
    "tdbg_column_stage_encoded": (C.c_int32, [C.c_void_p, C.POINTER(ColumnKey), C.c_uint32, C.c_uint32, C.c_void_p,
                                              C.c_uint64, C.c_uint64, C.c_uint32, C.POINTER(BlockStats)]),
    "tdbg_column_stage": (C.c_int32, [C.c_void_p, C.POINTER(ColumnKey), C.c_uint32, C.c_void_p, C.c_void_p,
                                      C.c_uint64, C.POINTER(BlockStats)]),
    "tdbg_column_evict": (C.c_int32, [C.c_void_p, C.POINTER(ColumnKey)]),
    "tdbg_segment_open": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "tdbg_segment_close": (None, [C.c_void_p]),
    "tdbg_segment_get_info": (C.c_int32, [C.c_void_p, C.POINTER(SegmentInfo)]),
    "tdbg_segment_block_info": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(SegmentBlock)]),
    "tdbg_segment_read_block": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64]),
    "tdbg_segment_stage": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
    "tdbg_table_reserve": (C.c_int32, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "tdbg_table_drop": (C.c_int32, [C.c_void_p, C.c_uint64]),
    "tdbg_plan_create": (C.c_int32, [C.c_void_p, C.POINTER(PlanDesc), C.POINTER(C.c_void_p)]),
    "tdbg_plan_destroy": (None, [C.c_void_p]),
    "tdbg_execute": (C.c_int32, [C.c_void_p, C.POINTER(RgRef), C.c_uint32, C.POINTER(C.c_void_p)]),
    "tdbg_execute_noalloc": (C.c_int32, [C.c_void_p, C.POINTER(RgRef), C.c_uint32, C.c_void_p, C.c_uint64,
                                         C.POINTER(ColumnView), C.c_uint32, C.POINTER(C.c_uint64)]),
    "tdbg_plan_last_stats": (C.c_int32, [C.c_void_p, C.POINTER(ExecStats)]),
    "tdbg_cancel": (C.c_int32, [C.c_void_p]),
    "tdbg_batch_create": (C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(C.c_void_p)]),
    "tdbg_batch_autotune": (C.c_int32, [C.c_void_p, C.c_uint32]),
    "tdbg_batch_execute": (C.c_int32, [C.c_void_p]),
    "tdbg_batch_result": (C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "tdbg_batch_last_stats": (C.c_int32, [C.c_void_p, C.POINTER(BatchStats)]),
    "tdbg_batch_plan_info": (C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(BatchRoute)]),
    "tdbg_batch_destroy": (None, [C.c_void_p]),
    "tdbg_result_num_rows": (C.c_uint64, [C.c_void_p]),
    "tdbg_result_num_columns": (C.c_uint32, [C.c_void_p]),
    "tdbg_result_column": (C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(ColumnView)]),
    "tdbg_result_free": (None, [C.c_void_p]),
}

_lib: Optional[C.CDLL] = None


def load(path: Optional[Path] = None) -> C.CDLL:
    """Loads the engine library and binds every export; raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else _LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built: run __graft_entry__.build() (the GPU engine has no fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.tdbg_abi_version() != ABI_VERSION:
        raise RuntimeError(f"ABI version {lib.tdbg_abi_version()} != {ABI_VERSION}")
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> int:
    if rc >= 0:
        return rc
    msg = load().tdbg_last_error().decode(errors="replace")
    st = Status(rc) if rc in Status._value2member_map_ else Status.DEVICE
    raise _ERRORS.get(st, GpuError)(f"{st.name}: {msg}")


# ---- objects -------------------------------------------------------------------

class Context:
    """One GPU's engine context (streams, HBM column store, admission budget)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        lib = load()
        h = C.c_void_p()
        check(lib.tdbg_context_create(device, hbm_budget_bytes, C.byref(h)))
        self._h = h
        self._lib = lib

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def bytes_in_use(self) -> int:
        return int(self._lib.tdbg_context_bytes_in_use(self._h))

    def stage(self, table_uid: int, segment_id: int, rg_idx: int, column_idx: int, type_: Type,
              values: np.ndarray, validity: Optional[np.ndarray] = None, stats=None) -> None:
        """Copies one decoded column block into HBM (cache.Key identity)."""
        values = np.ascontiguousarray(values, dtype=NUMPY_DTYPE[Type(type_)])
        key = ColumnKey(table_uid, segment_id, rg_idx, column_idx)
        vptr = None
        if validity is not None:
            validity = np.ascontiguousarray(validity, dtype=np.uint8)
            vptr = validity.ctypes.data_as(C.c_void_p)
        st = None
        if stats is not None:
            st = BlockStats()
            st.has_minmax = 1
            if Type(type_) in (Type.FLOAT, Type.DOUBLE):
                st.min_f, st.max_f = float(stats[0]), float(stats[1])
            else:
                st.min_i, st.max_i = int(stats[0]), int(stats[1])
        check(self._lib.tdbg_column_stage(self._h, C.byref(key), int(type_), values.ctypes.data_as(C.c_void_p),
                                          vptr, len(values), C.byref(st) if st is not None else None))

    def stage_encoded(self, table_uid: int, segment_id: int, rg_idx: int, column_idx: int, type_: Type,
                      encoding: Encoding, payload: bytes, n_rows: int, has_nulls: bool = False, stats=None) -> None:
        """Stages one block as thinDB's block cache holds it: the validity
        bitmap first when `has_nulls`, then the raw/FOR/RLE body; the
        library expands it in HBM."""
        key = ColumnKey(table_uid, segment_id, rg_idx, column_idx)
        buf = np.frombuffer(bytes(payload), dtype=np.uint8)
        st = None
        if stats is not None:
            st = BlockStats()
            st.has_minmax = 1
            st.min_i, st.max_i = int(stats[0]), int(stats[1])
        check(self._lib.tdbg_column_stage_encoded(self._h, C.byref(key), int(type_), int(encoding),
                                                  buf.ctypes.data_as(C.c_void_p), len(buf), n_rows,
                                                  1 if has_nulls else 0, C.byref(st) if st is not None else None))

    def stage_segment(self, segment: "Segment", table_uid: int, rg_idx: int, column_idx: int, type_: Type) -> None:
        """Reads one block of a segment file and stages it (cache.Key =
        {table_uid, the segment's id, rg_idx, column_idx})."""
        check(self._lib.tdbg_segment_stage(self._h, segment.handle, table_uid, rg_idx, column_idx, int(type_)))

    def evict(self, table_uid: int, segment_id: int, rg_idx: int, column_idx: int) -> None:
        key = ColumnKey(table_uid, segment_id, rg_idx, column_idx)
        check(self._lib.tdbg_column_evict(self._h, C.byref(key)))

    def reserve(self, table_uid: int, rows: int) -> None:
        check(self._lib.tdbg_table_reserve(self._h, table_uid, rows))

    def drop_table(self, table_uid: int) -> None:
        check(self._lib.tdbg_table_drop(self._h, table_uid))

    def close(self) -> None:
        if self._h:
            self._lib.tdbg_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ResultColumn:
    __slots__ = ("type", "scale", "nullable", "values", "validity")

    def __init__(self, type_: Type, scale: int, nullable: bool, values: np.ndarray, validity: Optional[np.ndarray]):
        self.type = type_
        self.scale = scale
        self.nullable = nullable
        self.values = values
        self.validity = validity  # packed bitmap (bit 1 = valid) or None

    def is_valid(self, row: int) -> bool:
        if self.validity is None:
            return True
        return bool((self.validity[row >> 3] >> (row & 7)) & 1)


def _copy_out(addr, n: int, dtype) -> np.ndarray:
    """n values of `dtype` copied out of host memory at `addr` (one memmove)."""
    a = np.empty(n, dtype)
    if n:
        C.memmove(a.ctypes.data, addr, a.nbytes)
    return a


def _column_from_view(v: ColumnView) -> ResultColumn:
    t = Type(v.type)
    n = int(v.n_rows)
    if t == Type.DECIMAL128:
        raw = _copy_out(v.data, 2 * n, np.uint64)
        lo = raw[0::2]
        hi = raw[1::2].view(np.int64)
        if np.array_equal(hi, lo.view(np.int64) >> 63):  # every value fits 64 bits
            vals = lo.view(np.int64).astype(object)
        else:
            vals = hi.astype(object) * (1 << 64) + lo.astype(object)
    else:
        vals = _copy_out(v.data, n, NUMPY_DTYPE[t])
    valid = None
    if v.validity:
        valid = _copy_out(v.validity, (n + 7) // 8, np.uint8)
    return ResultColumn(t, int(v.scale), bool(v.nullable), vals, valid)


def having_node(t: tuple) -> HavingNodeC:
    """One tdbg_having_node from (kind, ...) as Plan documents it."""
    kind = HavingKind(t[0])
    h = HavingNodeC()
    h.kind = int(kind)
    if kind == HavingKind.CMP:
        _, target, op, lit = t
        h.target = int(target)
        h.op = int(op)
        if isinstance(lit, float):
            h.lit_float = 1
            h.lit_f64 = lit
        else:
            v = int(lit)
            if not -(1 << 127) <= v < (1 << 127):
                raise Unsupported(f"HAVING literal {v} exceeds 128 bits")
            u = v & ((1 << 128) - 1)
            h.lit_lo = C.c_int64(u & ((1 << 64) - 1)).value
            h.lit_hi = C.c_int64(u >> 64).value
    elif kind in (HavingKind.IS_NULL, HavingKind.IS_NOT_NULL):
        h.target = int(t[1])
    elif kind in (HavingKind.AND, HavingKind.OR):
        h.arity = int(t[1])
    return h


def filter_node(t: tuple) -> FilterNodeC:
    """One tdbg_filter_node from (kind, ...) as Plan documents it."""
    kind = HavingKind(t[0])
    f = FilterNodeC()
    f.kind = int(kind)
    if kind == HavingKind.CMP:
        _, col, typ, op, lit = t
        f.column_idx, f.type, f.op = int(col), int(typ), int(op)
        if Type(typ) in (Type.FLOAT, Type.DOUBLE):
            f.f64 = float(lit)
        else:
            f.i64 = int(lit)
    elif kind in (HavingKind.IS_NULL, HavingKind.IS_NOT_NULL):
        f.column_idx = int(t[1])
    elif kind in (HavingKind.AND, HavingKind.OR):
        f.arity = int(t[1])
    return f


class Segment:
    """A thinDB segment file (.dat) parsed from a host image of the file
    (bytes, or an np.memmap of it); the image is kept alive with it."""

    def __init__(self, image, n_columns: int):
        lib = load()
        self._img = np.frombuffer(image, dtype=np.uint8) if not isinstance(image, np.ndarray) else image
        h = C.c_void_p()
        check(lib.tdbg_segment_open(self._img.ctypes.data_as(C.c_void_p), self._img.nbytes, n_columns, C.byref(h)))
        self._h = h
        self._lib = lib

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def info(self) -> SegmentInfo:
        out = SegmentInfo()
        check(self._lib.tdbg_segment_get_info(self._h, C.byref(out)))
        return out

    def block(self, rg_idx: int, column_idx: int) -> SegmentBlock:
        out = SegmentBlock()
        check(self._lib.tdbg_segment_block_info(self._h, rg_idx, column_idx, C.byref(out)))
        return out

    def read_block(self, rg_idx: int, column_idx: int) -> bytes:
        """The block's payload as thinDB's block cache holds it."""
        n = self.block(rg_idx, column_idx).payload_bytes
        buf = np.empty(max(n, 1), dtype=np.uint8)
        check(self._lib.tdbg_segment_read_block(self._h, rg_idx, column_idx, buf.ctypes.data_as(C.c_void_p), n))
        return buf[:n].tobytes()

    def close(self) -> None:
        if self._h:
            self._lib.tdbg_segment_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """A grouped query prepared on one context (tdbg_plan)."""

    def __init__(self, ctx: Context, table_uid: int, keys: Sequence[tuple], aggs: Sequence[tuple],
                 preds: Sequence[tuple] = (), order: Sequence[tuple] = (), limit: int = 0, offset: int = 0,
                 est_groups: int = 0, having: Sequence[tuple] = (), where: Sequence[tuple] = ()):
        """keys: (column_idx, Type); aggs: (AggOp, column_idx|-1, Type, scale, nullable);
        preds: (column_idx, Type, PredOp, literal, nullable); order: (target, desc);
        having: postfix nodes (HavingKind.CMP, target, PredOp, int|float literal),
        (IS_NULL|IS_NOT_NULL, target), (NOT,), (AND|OR, arity);
        where: postfix WHERE tree AND-ed with preds, (HavingKind.CMP, column_idx, Type, PredOp,
        literal), (IS_NULL|IS_NOT_NULL, column_idx), (NOT,), (AND|OR, arity)."""
        lib = load()
        self._lib = lib
        self.ctx = ctx
        self._keys = (KeyPart * max(1, len(keys)))(*[KeyPart(c, int(t)) for c, t in keys])
        self._aggs = (AggSpecC * max(1, len(aggs)))(
            *[AggSpecC(int(op), int(c), int(t), int(sc), int(bool(nl))) for op, c, t, sc, nl in aggs])
        pl = []
        for c, t, op, lit, nl in preds:
            p = PredicateC(c, int(t), int(op), int(bool(nl)))
            if Type(t) in (Type.FLOAT, Type.DOUBLE):
                p.f64 = float(lit)
            else:
                p.i64 = int(lit)
            pl.append(p)
        self._preds = (PredicateC * max(1, len(pl)))(*pl)
        self._order = (OrderSpecC * max(1, len(order)))(*[OrderSpecC(t, int(bool(d))) for t, d in order])
        self._having = (HavingNodeC * max(1, len(having)))(*[having_node(h) for h in having])
        self._where = (FilterNodeC * max(1, len(where)))(*[filter_node(w) for w in where])
        desc = PlanDesc(table_uid, len(keys), len(aggs), self._keys, self._aggs, len(pl), len(order), self._preds,
                        self._order, limit, offset, est_groups, len(having), len(where), self._having, self._where)
        h = C.c_void_p()
        check(lib.tdbg_plan_create(ctx.handle, C.byref(desc), C.byref(h)))
        self._h = h
        self.n_outputs = len(keys) + len(aggs)

    @staticmethod
    def rg_array(rgs: Sequence[tuple]):
        """A reusable C array of (segment_id, row_group_idx) references."""
        arr = (RgRef * max(1, len(rgs)))()
        for i, (s, r) in enumerate(rgs):
            arr[i].segment_id = s
            arr[i].row_group_idx = r
        return arr, len(rgs)

    def execute(self, rgs=None) -> list:
        """Runs the plan over `rgs` (None = every staged row group, a list of
        (segment_id, row_group_idx), or a prepared :meth:`rg_array`); returns
        the result columns (keys, then aggregates)."""
        out = C.c_void_p()
        if rgs is None:
            check(self._lib.tdbg_execute(self._h, None, 0, C.byref(out)))
        else:
            arr, n = rgs if isinstance(rgs, tuple) and len(rgs) == 2 and isinstance(rgs[1], int) \
                and not isinstance(rgs[0], int) else self.rg_array(rgs)
            check(self._lib.tdbg_execute(self._h, arr, n, C.byref(out)))
        return _result_columns(self._lib, out)

    def execute_noalloc(self, out_capacity: int, rgs=None) -> list:
        """tdbg_execute_noalloc: the same result as :meth:`execute`, with no
        allocation in the library (its scratch array and `out_capacity`
        bytes of host memory given here)."""
        buf = (C.c_uint8 * max(1, out_capacity))()
        views = (ColumnView * self.n_outputs)()
        n = C.c_uint64()
        arr, cnt = (None, 0) if rgs is None else self.rg_array(rgs)
        check(self._lib.tdbg_execute_noalloc(self._h, arr, cnt, buf, out_capacity, views, self.n_outputs,
                                             C.byref(n)))
        cols = [_column_from_view(views[i]) for i in range(self.n_outputs)]
        assert all(len(c.values) == n.value for c in cols)
        return cols

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def stats(self) -> ExecStats:
        st = ExecStats()
        check(self._lib.tdbg_plan_last_stats(self._h, C.byref(st)))
        return st

    def cancel(self) -> None:
        check(self._lib.tdbg_cancel(self._h))

    def close(self) -> None:
        if self._h:
            self._lib.tdbg_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _result_columns(lib, out) -> list:
    try:
        cols = []
        for i in range(lib.tdbg_result_num_columns(out)):
            v = ColumnView()
            check(lib.tdbg_result_column(out, i, C.byref(v)))
            cols.append(_column_from_view(v))
        return cols
    finally:
        lib.tdbg_result_free(out)


class Batch:
    """Many plans executed together in one persistent launch (tdbg_batch)."""

    STRATEGIES = {1: "lds_direct", 2: "hbm_atomics", 3: "lds_hash", 4: "radix"}

    def __init__(self, ctx: Context, plans: Sequence["Plan"]):
        lib = load()
        self._lib = lib
        self.ctx = ctx
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p.handle.value for p in self.plans])
        h = C.c_void_p()
        check(lib.tdbg_batch_create(ctx.handle, arr, len(self.plans), C.byref(h)))
        self._h = h

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def autotune(self, reps: int = 2) -> None:
        check(self._lib.tdbg_batch_autotune(self._h, reps))

    def execute(self) -> None:
        check(self._lib.tdbg_batch_execute(self._h))

    def result(self, i: int) -> list:
        """Result columns of plan i of the last execution (keys, then aggregates)."""
        out = C.c_void_p()
        check(self._lib.tdbg_batch_result(self._h, i, C.byref(out)))
        return _result_columns(self._lib, out)

    def stats(self) -> BatchStats:
        st = BatchStats()
        check(self._lib.tdbg_batch_last_stats(self._h, C.byref(st)))
        return st

    def route(self, i: int) -> BatchRoute:
        r = BatchRoute()
        check(self._lib.tdbg_batch_plan_info(self._h, i, C.byref(r)))
        return r

    def close(self) -> None:
        if self._h:
            self._lib.tdbg_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerConfig(C.Structure):
    _fields_ = [("lanes", C.c_uint32), ("max_batch", C.c_uint32), ("admit_bytes", C.c_uint64),
                ("large_rows", C.c_uint64)]


class TicketReport(C.Structure):
    _fields_ = [("submit_ns", C.c_int64), ("start_ns", C.c_int64), ("done_ns", C.c_int64), ("est_bytes", C.c_uint64),
                ("batch_size", C.c_uint32), ("lane", C.c_uint32), ("status", C.c_int32), ("pad", C.c_uint32)]


class ServerCounters(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("units", C.c_uint64),
                ("batched", C.c_uint64), ("peak_admitted", C.c_uint64), ("admission_waits", C.c_uint64)]


EXPORTS.update({
    "tdbg_server_start": (C.c_int32, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "tdbg_server_submit": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "tdbg_server_submit_rgs": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "tdbg_ticket_poll": (C.c_int32, [C.c_void_p]),
    "tdbg_ticket_wait": (C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "tdbg_ticket_wait_for": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "tdbg_ticket_info": (C.c_int32, [C.c_void_p, C.POINTER(TicketReport)]),
    "tdbg_ticket_free": (None, [C.c_void_p]),
    "tdbg_server_stats": (C.c_int32, [C.c_void_p, C.POINTER(ServerCounters)]),
    "tdbg_server_stop": (None, [C.c_void_p]),
})


class Ticket:
    """One submitted request (tdbg_ticket)."""

    def __init__(self, lib, h: C.c_void_p):
        self._lib = lib
        self._h = h

    def done(self) -> bool:
        rc = self._lib.tdbg_ticket_poll(self._h)
        if rc == Status.PENDING:
            return False
        check(rc)
        return True

    def wait(self) -> list:
        """Blocks until the request has run; returns its result columns."""
        out = C.c_void_p()
        check(self._lib.tdbg_ticket_wait(self._h, C.byref(out)))
        return _result_columns(self._lib, out)

    def wait_for(self, timeout_us: int):
        """Waits up to about `timeout_us`; the result columns, or None while
        the request is still queued or running."""
        out = C.c_void_p()
        rc = self._lib.tdbg_ticket_wait_for(self._h, timeout_us, C.byref(out))
        if rc == Status.PENDING:
            return None
        check(rc)
        return _result_columns(self._lib, out)

    def report(self) -> TicketReport:
        r = TicketReport()
        check(self._lib.tdbg_ticket_info(self._h, C.byref(r)))
        return r

    def close(self) -> None:
        if self._h:
            self._lib.tdbg_ticket_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The device request loop (tdbg_server): admission, continuous batching, lanes."""

    def __init__(self, ctx: Context, lanes: int = 0, max_batch: int = 0, admit_bytes: int = 0, large_rows: int = 0):
        lib = load()
        self._lib = lib
        self.ctx = ctx
        cfg = ServerConfig(lanes, max_batch, admit_bytes, large_rows)
        h = C.c_void_p()
        check(lib.tdbg_server_start(ctx.handle, C.byref(cfg), C.byref(h)))
        self._h = h

    def submit(self, plan: Plan, rgs=None) -> Ticket:
        """Queues one execution of `plan` over `rgs` (None = every staged row
        group, else a list of (segment_id, row_group_idx) or a prepared
        :meth:`Plan.rg_array`)."""
        t = C.c_void_p()
        if rgs is None:
            check(self._lib.tdbg_server_submit(self._h, plan.handle, C.byref(t)))
        else:
            arr, n = rgs if isinstance(rgs, tuple) and len(rgs) == 2 and isinstance(rgs[1], int) \
                and not isinstance(rgs[0], int) else Plan.rg_array(rgs)
            check(self._lib.tdbg_server_submit_rgs(self._h, plan.handle, C.cast(arr, C.c_void_p), n, C.byref(t)))
        return Ticket(self._lib, t)

    def counters(self) -> ServerCounters:
        c = ServerCounters()
        check(self._lib.tdbg_server_stats(self._h, C.byref(c)))
        return c

    def stop(self) -> None:
        if self._h:
            self._lib.tdbg_server_stop(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
