"""thinDB's grouped-operator interface over the GPU engine.

This mirrors the reference's operator surface for the hot path so a query
plan swaps the GPU operator in for the CPU one:

* :class:`Request` — src/exec/v2_shape_group_topn.zig:63-101 (``Request``),
  with :class:`AggSpec` (src/exec/aggregate.zig:237-268), :class:`Predicate`
  (src/exec/predicate.zig:35-51) and :class:`SortSpec`
  (src/exec/sort.zig:28-31).
* :func:`try_build_gpu_group` — the builder a planner calls at the seam
  src/exec/v2_pipeline.zig:17-34, before ``tryBuildLowCardGroup``
  (src/exec/engine_v2.zig:968); ``None`` declines to the CPU path.
* :class:`GpuGroupQuery` — the operator (src/exec/exec.zig:299-433 VTable:
  ``next``, ``deinit``, ``outputSchema``, ``addPrune``, ``explain``): the
  first :meth:`~GpuGroupQuery.next` runs the whole query and returns ONE
  :class:`Batch` (keys, then aggregates, ordered and sliced), later calls
  return ``None`` — as ``LowCardGroup.next`` (v2_lowcard_group.zig:1775-1784).
* :class:`Table` — the host view of a table's immutable segments and row
  groups (DESIGN.md of the reference :865-951), whose decoded blocks are
  staged into HBM once and reused across queries.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from decimal import Decimal
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import gpu
from .gpu import AggOp, PredOp, Type

ROW_GROUP_SIZE = 65536       # src/api/api.zig:140 default row_group_size
SEGMENT_ROWS = 1 << 20       # 1M-row segments


# ---- schema ---------------------------------------------------------------------

@dataclass(frozen=True)
class Column:
    """src/types.zig Column: name, type, nullability (+ DECIMAL scale)."""
    name: str
    type: Type
    nullable: bool = False
    scale: int = 0


@dataclass(frozen=True)
class AggSpec:
    """src/exec/aggregate.zig:237-268; func in count/count_distinct/sum/avg/min/max/udf."""
    func: str
    col: Optional[str]
    as_: str
    udf_name: Optional[str] = None


@dataclass(frozen=True)
class Predicate:
    """A ``col op literal`` leaf (src/exec/predicate.zig:35-51)."""
    col: str
    op: str  # eq, neq, lt, lte, gt, gte
    val: Union[int, float, Decimal]


@dataclass(frozen=True)
class And:
    children: Tuple["PredicateExpr", ...]


@dataclass(frozen=True)
class Or:
    children: Tuple["PredicateExpr", ...]


@dataclass(frozen=True)
class Not:
    child: "PredicateExpr"


@dataclass(frozen=True)
class IsNull:
    """``col IS NULL`` (PredicateExpr.is_null)."""
    col: str


@dataclass(frozen=True)
class IsNotNull:
    """``col IS NOT NULL`` (PredicateExpr.is_not_null)."""
    col: str


PredicateExpr = Union[Predicate, And, Or, Not, IsNull, IsNotNull]


@dataclass(frozen=True)
class SortSpec:
    col: str
    desc: bool = False


@dataclass
class Request:
    """src/exec/v2_shape_group_topn.zig:63-101 (the fields this path reads)."""
    group_cols: Sequence[str]
    aggs: Sequence[AggSpec]
    order_specs: Sequence[SortSpec] = ()
    limit: int = 0
    offset: int = 0
    where_filter: Optional[PredicateExpr] = None
    having_filter: Optional[PredicateExpr] = None
    derived: Sequence[object] = ()
    dop: int = 1
    udf_registry: Optional[object] = None
    est_groups: int = 0
    # the WHERE's conjuncts over derived / computed columns (:71-76): the GPU
    # operator declines a request that carries either
    residual_filter: Optional[PredicateExpr] = None
    computed_where: Optional[object] = None


# ---- tables -------------------------------------------------------------------------

@dataclass
class RowGroup:
    segment_id: int
    idx: int
    start: int      # first row in the table's arrays
    n_rows: int


class Table:
    """A table's columns as host arrays, cut into segments of ``segment_rows``
    rows and row groups of ``rg_rows`` rows.  Arrays are sliced, not copied."""

    _next_uid = 1

    def __init__(self, schema: Sequence[Column], columns: Dict[str, np.ndarray],
                 validity: Optional[Dict[str, np.ndarray]] = None, rg_rows: int = ROW_GROUP_SIZE,
                 segment_rows: int = SEGMENT_ROWS, uid: Optional[int] = None):
        self.schema = list(schema)
        self.columns = {c.name: np.ascontiguousarray(columns[c.name], dtype=gpu.NUMPY_DTYPE[c.type])
                        for c in self.schema}
        self.validity = dict(validity or {})  # name -> bool array (True = valid)
        for c in self.schema:
            if c.name in self.validity and not c.nullable:
                raise ValueError(f"column {c.name} is not nullable")
        n = {len(v) for v in self.columns.values()}
        if len(n) != 1:
            raise ValueError("columns differ in length")
        self.n_rows = n.pop()
        self.rg_rows = rg_rows
        self.segment_rows = segment_rows
        if uid is None:
            uid = Table._next_uid
            Table._next_uid += 1
        self.uid = uid
        self.row_groups: List[RowGroup] = []
        seg = 0
        for s0 in range(0, self.n_rows, segment_rows):
            s1 = min(self.n_rows, s0 + segment_rows)
            for i, r0 in enumerate(range(s0, s1, rg_rows)):
                self.row_groups.append(RowGroup(seg, i, r0, min(rg_rows, s1 - r0)))
            seg += 1
        self._packed: Dict[Tuple[str, int], np.ndarray] = {}

    def column(self, name: str) -> Column:
        for c in self.schema:
            if c.name.lower() == name.lower():
                return c
        raise KeyError(name)

    def column_index(self, name: str) -> int:
        for i, c in enumerate(self.schema):
            if c.name.lower() == name.lower():
                return i
        raise KeyError(name)

    def block(self, name: str, rg: RowGroup) -> Tuple[np.ndarray, Optional[np.ndarray]]:
        """One decoded block: values and a packed validity bitmap (bit 1 = valid)."""
        vals = self.columns[name][rg.start:rg.start + rg.n_rows]
        if name not in self.validity:
            return vals, None
        key = (name, rg.start)
        bm = self._packed.get(key)
        if bm is None:
            v = np.asarray(self.validity[name][rg.start:rg.start + rg.n_rows], dtype=bool)
            bm = np.packbits(v, bitorder="little")
            self._packed[key] = bm
        return vals, bm


# ---- batches ------------------------------------------------------------------------

@dataclass
class Batch:
    """src/exec/exec.zig:175-247: schema + column views + row count."""
    schema: List[Column]
    values: List[gpu.ResultColumn]
    row_count: int

    def column(self, name: str) -> gpu.ResultColumn:
        for c, v in zip(self.schema, self.values):
            if c.name.lower() == name.lower():
                return v
        raise KeyError(name)

    def rows(self) -> List[tuple]:
        """Rows as Python tuples (None for NULL)."""
        out = []
        for r in range(self.row_count):
            out.append(tuple((v.values[r].item() if hasattr(v.values[r], "item") else v.values[r])
                             if v.is_valid(r) else None for v in self.values))
        return out


# ---- request -> plan ------------------------------------------------------------------

_AGG_FUNCS = {"count", "count_distinct", "sum", "avg", "min", "max"}
_OPS = {"eq": PredOp.EQ, "neq": PredOp.NEQ, "lt": PredOp.LT, "lte": PredOp.LTE, "gt": PredOp.GT, "gte": PredOp.GTE}
_INT_TYPES = {Type.BOOLEAN, Type.TINYINT, Type.SMALLINT, Type.INT, Type.BIGINT, Type.DATE, Type.DATETIME,
              Type.DECIMAL64}
_FLOAT_TYPES = {Type.FLOAT, Type.DOUBLE}


class Declined(Exception):
    """The GPU operator does not cover this shape."""


def _conjuncts(expr: Optional[PredicateExpr]) -> list:
    """The top-level AND's children (nested ANDs flattened)."""
    if expr is None:
        return []
    if isinstance(expr, And):
        out: list = []
        for ch in expr.children:
            out.extend(_conjuncts(ch))
        return out
    return [expr]


def _where_postfix(expr, table: "Table", out: list, cols: List[str]) -> None:
    """A WHERE subtree in postfix order as gpu.Plan `where` nodes
    (PredicateExpr leaf / is_null / is_not_null / not / and / or,
    src/exec/predicate.zig:55-160); other forms decline."""
    if isinstance(expr, Predicate):
        c = table.column(expr.col)
        if c.type not in _INT_TYPES and c.type not in _FLOAT_TYPES:
            raise Declined(f"predicate on {expr.col}")
        op, lit = _coerce(c, expr.op, expr.val)
        out.append((gpu.HavingKind.CMP, table.column_index(expr.col), c.type, op, lit))
        cols.append(c.name)
    elif isinstance(expr, (IsNull, IsNotNull)):
        c = table.column(expr.col)
        kind = gpu.HavingKind.IS_NULL if isinstance(expr, IsNull) else gpu.HavingKind.IS_NOT_NULL
        out.append((kind, table.column_index(expr.col)))
        cols.append(c.name)
    elif isinstance(expr, Not):
        _where_postfix(expr.child, table, out, cols)
        out.append((gpu.HavingKind.NOT,))
    elif isinstance(expr, (And, Or)):
        for ch in expr.children:
            _where_postfix(ch, table, out, cols)
        out.append((gpu.HavingKind.AND if isinstance(expr, And) else gpu.HavingKind.OR, len(expr.children)))
    else:
        raise Declined(f"WHERE form {type(expr).__name__}")


def _coerce(col: Column, op: str, val) -> Tuple[PredOp, Union[int, float]]:
    """Coerces a literal to the column's physical domain (a DECIMAL literal to
    its scaled mantissa, a fractional literal on an integer column to the
    equivalent integer bound)."""
    if col.type in _FLOAT_TYPES:
        return _OPS[op], float(val)
    if isinstance(val, bool):
        val = int(val)
    x = Decimal(val) if not isinstance(val, float) else Decimal(repr(val))
    if col.type == Type.DECIMAL64:
        x = x.scaleb(col.scale)
    if x == x.to_integral_value():
        return _OPS[op], int(x)
    if op == "lt" or op == "gte":
        return _OPS[op], int(math.ceil(x))
    if op == "lte" or op == "gt":
        return _OPS[op], int(math.floor(x))
    raise Declined("fractional equality literal on an integer column")


_DECIMALS = (Type.DECIMAL64, Type.DECIMAL128)


def _having_number(val, col: Column) -> Tuple[bool, Union[int, float]]:
    """A HAVING literal in the output column's domain (valueAsNum,
    src/exec/v2_shape_group_topn.zig:2391-2407, after coercion): a DECIMAL
    output compares mantissas, so the literal is scaled by its scale; an
    integral result is an integer, anything else a double."""
    if isinstance(val, bool):
        val = int(val)
    if col.type not in _DECIMALS:
        if isinstance(val, float):
            return True, val
        if isinstance(val, int):
            return False, val
    x = Decimal(repr(val)) if isinstance(val, float) else Decimal(val)
    if col.type in _DECIMALS:
        x = x.scaleb(col.scale)
    if not x.is_finite():
        raise Declined("non-finite HAVING literal")
    if x == x.to_integral_value():
        return False, int(x)
    return True, float(x)


def _having_postfix(expr, schema: List[Column], out: list) -> None:
    """Appends the HAVING tree in postfix order as gpu.Plan nodes; declines
    the forms havingExprSupported (v2_shape_group_topn.zig:2443-2459)
    rejects."""
    names = [c.name.lower() for c in schema]

    def target(name: str) -> int:
        try:
            return names.index(name.lower())
        except ValueError:
            raise Declined(f"HAVING on {name}, which is not an output column") from None

    if isinstance(expr, Predicate):
        if expr.op not in _OPS:
            raise Declined(f"HAVING op {expr.op}")
        t = target(expr.col)
        is_f, v = _having_number(expr.val, schema[t])
        out.append((gpu.HavingKind.CMP, t, _OPS[expr.op], float(v) if is_f else int(v)))
    elif isinstance(expr, (IsNull, IsNotNull)):
        kind = gpu.HavingKind.IS_NULL if isinstance(expr, IsNull) else gpu.HavingKind.IS_NOT_NULL
        out.append((kind, target(expr.col)))
    elif isinstance(expr, Not):
        _having_postfix(expr.child, schema, out)
        out.append((gpu.HavingKind.NOT,))
    elif isinstance(expr, (And, Or)):
        for ch in expr.children:
            _having_postfix(ch, schema, out)
        out.append((gpu.HavingKind.AND if isinstance(expr, And) else gpu.HavingKind.OR, len(expr.children)))
    else:
        raise Declined(f"HAVING form {type(expr).__name__}")


def output_schema(table: Table, request: Request) -> List[Column]:
    """Keys then aggregates (v2_lowcard_group.zig:1674-1684; output types of
    src/exec/aggregate.zig:3700-3735)."""
    out = [Column(table.column(k).name, table.column(k).type, table.column(k).nullable, table.column(k).scale)
           for k in request.group_cols]
    for a in request.aggs:
        f = a.func.lower()
        if f in ("count", "count_distinct"):
            out.append(Column(a.as_, Type.BIGINT, False))
            continue
        c = table.column(a.col)
        if f == "sum":
            if c.type in _FLOAT_TYPES:
                out.append(Column(a.as_, Type.DOUBLE, True))
            elif c.type == Type.DECIMAL64:
                out.append(Column(a.as_, Type.DECIMAL128, True, c.scale))
            else:
                out.append(Column(a.as_, Type.BIGINT, True))
        elif f == "avg":
            out.append(Column(a.as_, Type.DOUBLE, True))
        else:
            out.append(Column(a.as_, c.type, True, c.scale))
    return out


class GpuGroupQuery:
    """The GPU grouped operator (exec.Query contract)."""

    def __init__(self, session: "GpuSession", table: Table, request: Request, plan: gpu.Plan,
                 columns: List[str], schema: List[Column]):
        self.session = session
        self.table = table
        self.request = request
        self.plan = plan
        self._columns = columns
        self._schema = schema
        self._done = False
        self._batch: Optional[Batch] = None

    def output_schema(self) -> List[Column]:
        return self._schema

    def add_prune(self, pred) -> None:  # the GPU scan prunes by zone maps itself
        return None

    def explain(self) -> str:
        r = self.request
        return (f"GPU group-by [{', '.join(r.group_cols)}] aggs={len(r.aggs)} "
                f"order={len(r.order_specs)} limit={r.limit} offset={r.offset}")

    def stats(self) -> dict:
        return {"upper_rows": self.request.est_groups or None}

    def next(self) -> Optional[Batch]:
        if self._done:
            return None
        self._done = True
        self.session.stage_table(self.table, self._columns)
        cols = self.plan.execute(self.plan.rg_array([(rg.segment_id, rg.idx) for rg in self.table.row_groups]))
        n = len(cols[0].values) if cols else 0
        self._batch = Batch(self._schema, cols, n)
        return self._batch

    def cancel(self) -> None:
        self.plan.cancel()

    def deinit(self) -> None:
        self.plan.close()
        self._batch = None


class GpuSession:
    """A GPU context plus the registry of staged blocks (what the HBM column
    cache holds), shared by every query of a process."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.ctx = gpu.Context(device, hbm_budget_bytes)
        self._staged: set = set()

    def stage_table(self, table: Table, columns: Sequence[str]) -> int:
        """Stages every block of `columns` not yet resident; returns bytes copied."""
        copied = 0
        need = [(n, table.column_index(n), table.column(n)) for n in dict.fromkeys(columns)]
        first = not any(k[0] == table.uid for k in self._staged)
        if first:
            self.ctx.reserve(table.uid, table.n_rows + len(table.row_groups) * 256)
        for rg in table.row_groups:
            for name, ci, col in need:
                key = (table.uid, rg.segment_id, rg.idx, ci)
                if key in self._staged:
                    continue
                vals, bm = table.block(col.name, rg)
                self.ctx.stage(table.uid, rg.segment_id, rg.idx, ci, col.type, vals, bm)
                self._staged.add(key)
                copied += vals.nbytes
        return copied

    def stage_segment(self, table: Table, segment: "gpu.Segment", columns: Optional[Sequence[str]] = None) -> int:
        """Stages the blocks of `columns` (default: every column) of one
        segment file of `table` straight from the file image: each block is
        decompressed and restored on the host and expanded in HBM.  Returns
        the number of blocks staged."""
        info = segment.info()
        names = list(columns) if columns is not None else [c.name for c in table.schema]
        if not any(k[0] == table.uid for k in self._staged):
            self.ctx.reserve(table.uid, table.n_rows + len(table.row_groups) * 256)
        n = 0
        for rg in range(info.n_row_groups):
            for name in names:
                ci = table.column_index(name)
                key = (table.uid, info.segment_id, rg, ci)
                if key in self._staged:
                    continue
                self.ctx.stage_segment(segment, table.uid, rg, ci, table.column(name).type)
                self._staged.add(key)
                n += 1
        return n

    def evict_table(self, table: Table) -> None:
        self.ctx.drop_table(table.uid)
        self._staged = {k for k in self._staged if k[0] != table.uid}

    def close(self) -> None:
        self.ctx.close()


def build_plan(session: GpuSession, table: Table, request: Request) -> Tuple[gpu.Plan, List[str]]:
    """Translates a Request into a tdbg plan; raises Declined for uncovered shapes."""
    if request.derived:
        raise Declined("derived columns stay on the CPU path")
    if request.residual_filter is not None or request.computed_where is not None:
        raise Declined("a WHERE over computed columns stays on the CPU path")
    if not request.group_cols or not request.aggs:
        raise Declined("needs group keys and aggregates")
    cols: List[str] = []
    keys = []
    for k in request.group_cols:
        c = table.column(k)
        if c.type not in _INT_TYPES:
            raise Declined(f"group key {k} is not integer-typed")
        keys.append((table.column_index(k), c.type))
        cols.append(c.name)
    aggs = []
    for a in request.aggs:
        f = a.func.lower()
        if f not in _AGG_FUNCS or a.udf_name:
            raise Declined(f"aggregate {a.func}")
        if f == "count" and a.col is None:
            aggs.append((AggOp.COUNT_STAR, -1, Type.BIGINT, 0, False))
            continue
        if a.col is None:
            raise Declined(f"{a.func} needs a column")
        c = table.column(a.col)
        if c.type not in _INT_TYPES and c.type not in _FLOAT_TYPES:
            raise Declined(f"aggregate input {a.col}")
        op = {"count": AggOp.COUNT, "count_distinct": AggOp.COUNT_DISTINCT, "sum": AggOp.SUM, "avg": AggOp.AVG,
              "min": AggOp.MIN, "max": AggOp.MAX}[f]
        aggs.append((op, table.column_index(a.col), c.type, c.scale, c.nullable))
        cols.append(c.name)
    # the WHERE's top-level `col op literal` conjuncts prune by zone maps and
    # fuse into the scan; the rest is one tree AND-ed with them
    preds = []
    where: list = []
    rest = []
    for p in _conjuncts(request.where_filter):
        if not isinstance(p, Predicate):
            rest.append(p)
            continue
        c = table.column(p.col)
        if c.type not in _INT_TYPES and c.type not in _FLOAT_TYPES:
            raise Declined(f"predicate on {p.col}")
        op, lit = _coerce(c, p.op, p.val)
        preds.append((table.column_index(p.col), c.type, op, lit, c.nullable))
        cols.append(c.name)
    for r in rest:
        _where_postfix(r, table, where, cols)
    if len(rest) > 1:
        where.append((gpu.HavingKind.AND, len(rest)))
    if len(where) > gpu.MAX_WHERE:
        raise Declined(f"WHERE tree of {len(where)} nodes")
    schema = output_schema(table, request)
    order = []
    for s in request.order_specs:
        idx = next((i for i, c in enumerate(schema) if c.name.lower() == s.col.lower()), None)
        if idx is None:
            raise Declined(f"ORDER BY {s.col} is not an output column")
        order.append((idx, s.desc))
    having: list = []
    if request.having_filter is not None:
        _having_postfix(request.having_filter, schema, having)
        if len(having) > gpu.MAX_HAVING:
            raise Declined(f"HAVING of {len(having)} nodes")
    try:
        plan = gpu.Plan(session.ctx, table.uid, keys, aggs, preds, order, request.limit, request.offset,
                        request.est_groups, having, where)
    except gpu.Unsupported as e:
        raise Declined(str(e)) from e
    return plan, cols


def try_build_gpu_group(session: GpuSession, table: Table, request: Request) -> Optional[GpuGroupQuery]:
    """The v2_pipeline seam: a GPU operator for this request, or None to decline."""
    try:
        plan, cols = build_plan(session, table, request)
    except Declined:
        return None
    return GpuGroupQuery(session, table, request, plan, cols, output_schema(table, request))


class GpuBatchRun:
    """Many grouped requests over (possibly different) tables run together in
    one persistent device launch (gpu.Batch); each result equals what the
    request returns alone.  ORDER BY / LIMIT / OFFSET and COUNT(DISTINCT)
    requests are declined: run those through try_build_gpu_group."""

    def __init__(self, session: GpuSession, pairs: Sequence[Tuple[Table, Request]]):
        self.session = session
        self.items = []
        plans = []
        try:
            for t, r in pairs:
                if r.order_specs or r.limit or r.offset:
                    raise Declined("ORDER BY / LIMIT / OFFSET requests run alone")
                if any(a.func.lower() == "count_distinct" for a in r.aggs):
                    raise Declined("COUNT(DISTINCT) requests run alone")
                plan, cols = build_plan(session, t, r)
                self.items.append((t, r, plan, output_schema(t, r)))
                plans.append(plan)
                session.stage_table(t, cols)
            self.batch = gpu.Batch(session.ctx, plans)
        except Exception:
            for p in plans:
                p.close()
            raise

    def autotune(self, reps: int = 2) -> None:
        self.batch.autotune(reps)

    def execute(self) -> List[Batch]:
        self.batch.execute()
        return [self.result(i) for i in range(len(self.items))]

    def result(self, i: int) -> Batch:
        cols = self.batch.result(i)
        return Batch(self.items[i][3], cols, len(cols[0].values) if cols else 0)

    def close(self) -> None:
        self.batch.close()
        for _, _, p, _ in self.items:
            p.close()
