"""Python binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  It translates a
``thindb_amd.exec.Request`` over a ``thindb_amd.exec.Table`` into the
oracle's plan with its own code (independent of the product's translation)
and returns the reference-semantics result.
"""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from decimal import Decimal
from pathlib import Path
from typing import List, Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

# enum tdbg_type / tdbg_agg_op / tdbg_pred_op values (include/thindb_gpu.h)
T_BOOLEAN, T_TINYINT, T_SMALLINT, T_INT, T_BIGINT, T_DATE, T_DATETIME, T_DECIMAL64, T_FLOAT, T_DOUBLE, \
    T_DECIMAL128 = range(1, 12)
FLOATS = (T_FLOAT, T_DOUBLE)
DTYPE = {T_BOOLEAN: np.uint8, T_TINYINT: np.int8, T_SMALLINT: np.int16, T_INT: np.int32, T_BIGINT: np.int64,
         T_DATE: np.int32, T_DATETIME: np.int64, T_DECIMAL64: np.int64, T_FLOAT: np.float32, T_DOUBLE: np.float64}
OPS = {"eq": 0, "neq": 1, "lt": 2, "lte": 3, "gt": 4, "gte": 5}
AGG = {"count*": 0, "count": 1, "sum": 2, "avg": 3, "min": 4, "max": 5, "count_distinct": 6}


class _Col(C.Structure):
    _fields_ = [("type", C.c_uint32), ("pad", C.c_uint32), ("data", C.c_void_p), ("validity", C.c_void_p)]


class _Table(C.Structure):
    _fields_ = [("n_rows", C.c_uint64), ("rg_rows", C.c_uint32), ("n_cols", C.c_uint32), ("cols", C.POINTER(_Col))]


class _Key(C.Structure):
    _fields_ = [("column_idx", C.c_uint32), ("type", C.c_uint32)]


class _Agg(C.Structure):
    _fields_ = [("op", C.c_uint32), ("column_idx", C.c_int32), ("input_type", C.c_uint32),
                ("input_scale", C.c_uint8), ("nullable", C.c_uint8), ("pad", C.c_uint8 * 2)]


class _Pred(C.Structure):
    _fields_ = [("column_idx", C.c_uint32), ("type", C.c_uint32), ("op", C.c_uint32), ("nullable", C.c_uint32),
                ("i64", C.c_int64), ("f64", C.c_double)]


class _Order(C.Structure):
    _fields_ = [("target", C.c_uint32), ("desc", C.c_uint32)]


class _Having(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("target", C.c_uint32), ("op", C.c_uint32), ("arity", C.c_uint32),
                ("lit_float", C.c_uint32), ("pad", C.c_uint32), ("lit_lo", C.c_uint64), ("lit_hi", C.c_uint64),
                ("lit_f64", C.c_double)]


class _Filter(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("column_idx", C.c_uint32), ("type", C.c_uint32), ("op", C.c_uint32),
                ("arity", C.c_uint32), ("pad", C.c_uint32), ("i64", C.c_int64), ("f64", C.c_double)]


class _Desc(C.Structure):
    _fields_ = [("table_uid", C.c_uint64), ("n_keys", C.c_uint32), ("n_aggs", C.c_uint32),
                ("keys", C.POINTER(_Key)), ("aggs", C.POINTER(_Agg)), ("n_preds", C.c_uint32),
                ("n_order", C.c_uint32), ("preds", C.POINTER(_Pred)), ("order", C.POINTER(_Order)),
                ("limit", C.c_uint64), ("offset", C.c_uint64), ("est_groups", C.c_uint64),
                ("n_having", C.c_uint32), ("n_where", C.c_uint32), ("having", C.POINTER(_Having)),
                ("where", C.POINTER(_Filter))]


H_CMP, H_IS_NULL, H_IS_NOT_NULL, H_NOT, H_AND, H_OR = range(6)


class _View(C.Structure):
    _fields_ = [("type", C.c_uint32), ("scale", C.c_uint32), ("nullable", C.c_uint32), ("pad", C.c_uint32),
                ("data", C.c_void_p), ("validity", C.c_void_p), ("n_rows", C.c_uint64)]


_lib = None


def build() -> Path:
    """Compiles oracle.c into liboracle.so (gcc, host only)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-pthread", str(HERE / "oracle.c"), "-o", str(tmp), "-lm"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode:
        raise RuntimeError(p.stderr)
    tmp.replace(LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.oracle_group_by.argtypes = [C.POINTER(_Desc), C.POINTER(_Table), C.c_uint32, C.POINTER(C.c_void_p)]
        L.oracle_group_by.restype = C.c_int32
        L.oracle_group_by_radix.argtypes = [C.POINTER(_Desc), C.POINTER(_Table), C.c_uint32, C.POINTER(C.c_void_p)]
        L.oracle_group_by_radix.restype = C.c_int32
        L.oracle_result_num_rows.argtypes = [C.c_void_p]
        L.oracle_result_num_rows.restype = C.c_uint64
        L.oracle_result_num_groups.argtypes = [C.c_void_p]
        L.oracle_result_num_groups.restype = C.c_uint64
        L.oracle_result_num_columns.argtypes = [C.c_void_p]
        L.oracle_result_num_columns.restype = C.c_uint32
        L.oracle_result_column.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(_View)]
        L.oracle_result_column.restype = C.c_int32
        L.oracle_result_compensated.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_double)]
        L.oracle_result_compensated.restype = C.c_int32
        L.oracle_result_free.argtypes = [C.c_void_p]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_plant.argtypes = [C.c_int]
        L.oracle_plant.restype = None
        _lib = L
    return _lib


class Unsupported(Exception):
    pass


class OracleColumn:
    def __init__(self, type_, scale, values, valid, compensated):
        self.type = type_
        self.scale = scale
        self.values = values          # numpy array (object for DECIMAL128)
        self.valid = valid            # bool array
        self.compensated = compensated  # float array for DOUBLE SUM/AVG, else None


class OracleResult:
    def __init__(self, columns: List[OracleColumn], n_groups: int):
        self.columns = columns
        self.n_groups = n_groups

    @property
    def n_rows(self) -> int:
        return len(self.columns[0].values) if self.columns else 0


def _literal(col_type: int, scale: int, op: str, val):
    """Coerces a literal to the column domain (own restatement)."""
    if col_type in FLOATS:
        return OPS[op], 0, float(val)
    x = Decimal(repr(val)) if isinstance(val, float) else Decimal(int(val) if isinstance(val, bool) else val)
    if col_type == T_DECIMAL64:
        x = x * (Decimal(10) ** scale)
    if x == x.to_integral_value():
        return OPS[op], int(x), 0.0
    if op in ("lt", "gte"):
        return OPS[op], int(math.ceil(x)), 0.0
    if op in ("lte", "gt"):
        return OPS[op], int(math.floor(x)), 0.0
    raise Unsupported("fractional equality literal")


def _flatten(expr) -> list:
    """The top-level AND's members (nested ANDs flattened)."""
    if expr is None:
        return []
    if type(expr).__name__ == "And":
        out = []
        for ch in expr.children:
            out.extend(_flatten(ch))
        return out
    return [expr]


def _where_nodes(expr, schema, idx) -> list:
    """A WHERE subtree in postfix order (children first)."""
    kind = type(expr).__name__
    if kind == "Predicate":
        i = idx[expr.col.lower()]
        c = schema[i]
        op, iv, fv = _literal(int(c.type), c.scale, expr.op, expr.val)
        return [_Filter(H_CMP, i, int(c.type), op, 0, 0, iv, fv)]
    if kind in ("IsNull", "IsNotNull"):
        return [_Filter(H_IS_NULL if kind == "IsNull" else H_IS_NOT_NULL, idx[expr.col.lower()])]
    if kind == "Not":
        return _where_nodes(expr.child, schema, idx) + [_Filter(H_NOT)]
    if kind in ("And", "Or"):
        out = []
        for ch in expr.children:
            out.extend(_where_nodes(ch, schema, idx))
        return out + [_Filter(H_AND if kind == "And" else H_OR, 0, 0, 0, len(expr.children))]
    raise Unsupported(f"WHERE form {kind}")


def _having_literal(val, out_type: int, out_scale: int):
    """A HAVING literal as a number in the output column's domain: scaled to
    the mantissa of a DECIMAL output; an integral value is an integer, any
    other a double.  Returns (is_float, int_value, float_value)."""
    if isinstance(val, bool):
        val = int(val)
    if isinstance(val, float) and out_type not in (T_DECIMAL64, T_DECIMAL128):
        return 1, 0, val
    x = Decimal(repr(val)) if isinstance(val, float) else Decimal(val)
    if out_type in (T_DECIMAL64, T_DECIMAL128):
        x = x * (Decimal(10) ** out_scale)
    if x.is_finite() and x == x.to_integral_value():
        return 0, int(x), 0.0
    return 1, 0, float(x)


def _having_nodes(expr, names: list, types: list, scales: list) -> list:
    """The predicate tree in postfix order (children first)."""
    kind = type(expr).__name__
    if kind == "Predicate":
        t = names.index(expr.col.lower())
        fl, iv, fv = _having_literal(expr.val, types[t], scales[t])
        h = _Having(H_CMP, t, OPS[expr.op], 0, fl)
        u = iv & ((1 << 128) - 1)
        h.lit_lo, h.lit_hi, h.lit_f64 = u & ((1 << 64) - 1), u >> 64, fv
        return [h]
    if kind in ("IsNull", "IsNotNull"):
        return [_Having(H_IS_NULL if kind == "IsNull" else H_IS_NOT_NULL, names.index(expr.col.lower()))]
    if kind == "Not":
        return _having_nodes(expr.child, names, types, scales) + [_Having(H_NOT)]
    if kind in ("And", "Or"):
        out = []
        for ch in expr.children:
            out.extend(_having_nodes(ch, names, types, scales))
        return out + [_Having(H_AND if kind == "And" else H_OR, 0, 0, len(expr.children))]
    raise Unsupported(f"HAVING form {kind}")


def _out_types(schema, idx, request) -> tuple:
    """Output column types and DECIMAL scales (keys, then aggregates)."""
    types, scales = [], []
    for k in request.group_cols:
        c = schema[idx[k.lower()]]
        types.append(int(c.type))
        scales.append(c.scale)
    for a in request.aggs:
        f = a.func.lower()
        if f in ("count", "count_distinct"):
            types.append(T_BIGINT)
            scales.append(0)
            continue
        c = schema[idx[a.col.lower()]]
        ct = int(c.type)
        if f == "avg" or (f == "sum" and ct in FLOATS):
            types.append(T_DOUBLE)
            scales.append(0)
        elif f == "sum":
            types.append(T_DECIMAL128 if ct == T_DECIMAL64 else T_BIGINT)
            scales.append(c.scale if ct == T_DECIMAL64 else 0)
        else:
            types.append(ct)
            scales.append(c.scale)
    return types, scales


def run(table, request, dop: int = 1, radix: bool = False) -> OracleResult:
    """Runs `request` over `table` with the CPU restatement (radix: through
    the silo's high-cardinality arrangement, oracle_group_by_radix: the same
    result, the CPU baseline for high-cardinality queries)."""
    L = lib()
    schema = table.schema
    idx = {c.name.lower(): i for i, c in enumerate(schema)}
    cols = (_Col * len(schema))()
    keep = []
    for i, c in enumerate(schema):
        arr = np.ascontiguousarray(table.columns[c.name], dtype=DTYPE[int(c.type)])
        keep.append(arr)
        cols[i].type = int(c.type)
        cols[i].data = arr.ctypes.data_as(C.c_void_p)
        if c.name in table.validity:
            bm = np.packbits(np.asarray(table.validity[c.name], dtype=bool), bitorder="little")
            keep.append(bm)
            cols[i].validity = bm.ctypes.data_as(C.c_void_p)
    t = _Table(table.n_rows, table.rg_rows, len(schema), cols)
    keys = (_Key * max(1, len(request.group_cols)))()
    for i, k in enumerate(request.group_cols):
        c = schema[idx[k.lower()]]
        keys[i].column_idx = idx[k.lower()]
        keys[i].type = int(c.type)
    aggs = (_Agg * max(1, len(request.aggs)))()
    out_names = [schema[idx[k.lower()]].name for k in request.group_cols]
    for i, a in enumerate(request.aggs):
        f = a.func.lower()
        if f == "count" and a.col is None:
            aggs[i].op = AGG["count*"]
            aggs[i].column_idx = -1
            aggs[i].input_type = T_BIGINT
        else:
            c = schema[idx[a.col.lower()]]
            aggs[i].op = AGG[f]
            aggs[i].column_idx = idx[a.col.lower()]
            aggs[i].input_type = int(c.type)
            aggs[i].input_scale = c.scale
            aggs[i].nullable = int(c.nullable)
        out_names.append(a.as_)
    members = _flatten(request.where_filter)
    leaves = [m for m in members if type(m).__name__ == "Predicate"]
    rest = [m for m in members if type(m).__name__ != "Predicate"]
    wl = []
    for m in rest:
        wl.extend(_where_nodes(m, schema, idx))
    if len(rest) > 1:
        wl.append(_Filter(H_AND, 0, 0, 0, len(rest)))
    where = (_Filter * max(1, len(wl)))(*wl)
    preds = (_Pred * max(1, len(leaves)))()
    for i, p in enumerate(leaves):
        c = schema[idx[p.col.lower()]]
        op, iv, fv = _literal(int(c.type), c.scale, p.op, p.val)
        preds[i].column_idx = idx[p.col.lower()]
        preds[i].type = int(c.type)
        preds[i].op = op
        preds[i].nullable = int(c.nullable)
        preds[i].i64 = iv
        preds[i].f64 = fv
    order = (_Order * max(1, len(request.order_specs)))()
    lower = [n.lower() for n in out_names]
    for i, s in enumerate(request.order_specs):
        order[i].target = lower.index(s.col.lower())
        order[i].desc = int(bool(s.desc))
    hl = []
    if getattr(request, "having_filter", None) is not None:
        types, scales = _out_types(schema, idx, request)
        hl = _having_nodes(request.having_filter, lower, types, scales)
    having = (_Having * max(1, len(hl)))(*hl)
    desc = _Desc(table.uid, len(request.group_cols), len(request.aggs), keys, aggs, len(leaves),
                 len(request.order_specs), preds, order, request.limit, request.offset, request.est_groups,
                 len(hl), len(wl), having, where)
    h = C.c_void_p()
    rc = (L.oracle_group_by_radix if radix else L.oracle_group_by)(C.byref(desc), C.byref(t), dop, C.byref(h))
    if rc == -2:
        raise Unsupported(L.oracle_last_error().decode())
    if rc:
        raise RuntimeError(f"oracle failed {rc}: {L.oracle_last_error().decode()}")
    try:
        n = L.oracle_result_num_rows(h)
        out = []
        for i in range(L.oracle_result_num_columns(h)):
            v = _View()
            L.oracle_result_column(h, i, C.byref(v))
            ty = int(v.type)
            if ty == T_DECIMAL128:
                raw = np.ctypeslib.as_array(C.cast(v.data, C.POINTER(C.c_uint64)), shape=(2 * n,)).copy() \
                    if n else np.zeros(0, np.uint64)
                vals = np.array([int(np.int64(raw[2 * r + 1].view(np.int64))) * (1 << 64) + int(raw[2 * r])
                                 for r in range(n)], dtype=object)
            elif n:
                ct = np.ctypeslib.as_ctypes_type(DTYPE[ty])
                vals = np.ctypeslib.as_array(C.cast(v.data, C.POINTER(ct)), shape=(n,)).copy()
            else:
                vals = np.zeros(0, DTYPE[ty])
            if v.validity:
                bm = np.ctypeslib.as_array(C.cast(v.validity, C.POINTER(C.c_uint8)), shape=((n + 7) // 8,))
                valid = np.unpackbits(bm, bitorder="little")[:n].astype(bool)
            else:
                valid = np.ones(n, bool)
            comp = None
            if ty == T_DOUBLE and n:
                buf = np.zeros(n, np.float64)
                if L.oracle_result_compensated(h, i, buf.ctypes.data_as(C.POINTER(C.c_double))) == 0:
                    comp = buf
            out.append(OracleColumn(ty, int(v.scale), vals, valid, comp))
        return OracleResult(out, int(L.oracle_result_num_groups(h)))
    finally:
        L.oracle_result_free(h)


class planted:
    """Context manager: the oracle's ORDER BY comparator with one known
    mistake planted (tests only; see oracle_plant in oracle.c)."""
    DECIMAL_LOW_WORD, AVG_TRUNC, NULL_AS_ZERO, TIES_BY_GID, NAN_FIRST, TIES_SIGNED = range(1, 7)

    def __init__(self, which: int):
        self.which = which

    def __enter__(self):
        lib().oracle_plant(self.which)
        return self

    def __exit__(self, *exc):
        lib().oracle_plant(0)
        return False
