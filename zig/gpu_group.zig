//! GPU grouped operator for thinDB: scan -> filter -> GROUP BY -> ORDER BY /
//! LIMIT on an MI355X through libthindb_gpu.so (include/thindb_gpu.h).
//!
//! A maintainer adds this file as src/exec/gpu_group.zig, a sibling builder
//! at the pipeline seam (src/exec/v2_pipeline.zig:1-34), and tries it before
//! tryBuildLowCardGroup in engine_v2.buildGroupTopN (src/exec/engine_v2.zig:
//! 968-972); see the end of this file.  tryBuild returns null to decline, and
//! the CPU cores run unchanged.
//!
//! Written against the reference's interfaces as they stand (no Zig toolchain
//! exists in this build's image, so it is not compiled here; the same C-ABI
//! calls are exercised by the Python tests through ctypes):
//!   - exec.Query / makeQuery / Batch        src/exec/exec.zig:175-247, :299-433, :1221-1239
//!   - explain(ptr, out, allocator, depth)   src/exec/exec.zig:409 (VTable.explain)
//!   - memory.checkCancelled(allocator)      src/memory.zig:60
//!   - Request (incl. residual_filter, computed_where)  src/exec/v2_shape_group_topn.zig:63-101
//!   - Scan.captureSnapshotAlloc / Snapshot  src/exec/scan.zig:1097-1137, :1206-1215
//!   - Table.segments_dir / segmentFileName  src/api/table.zig:86, :1255
//!   - ColumnStore.initCapacity / view / appendValidBit   src/engine/store.zig:247-361
//!   - aggregate.aggOutputTypeFor            src/exec/aggregate.zig:3618
const std = @import("std");
const Allocator = std.mem.Allocator;

const api = @import("../api/api.zig");
const types = @import("../types.zig");
const Column = types.Column;
const Type = types.Type;
const storage = @import("../storage/storage.zig");
const ColumnView = storage.ColumnView;
const store = @import("../engine/store.zig");
const ColumnStore = store.ColumnStore;
const exec = @import("exec.zig");
const aggregate = @import("aggregate.zig");
const Scan = @import("scan.zig").Scan;
const v2_pipeline = @import("v2_pipeline.zig");
const gpu_serving = @import("gpu_serving.zig");
const PredicateExpr = exec.PredicateExpr;

const c = @cImport(@cInclude("thindb_gpu.h"));

pub const Request = @import("v2_shape_group_topn.zig").Request;

// ---- one process-wide context per GPU (as core_scheduler is process-wide) ----

var g_ctx: ?*c.tdbg_context = null;
var g_ctx_once = std.once(initContext);

fn initContext() void {
    if (c.tdbg_abi_version() != c.TDBG_ABI_VERSION) return;
    var ctx: ?*c.tdbg_context = null;
    if (c.tdbg_context_create(0, 0, &ctx) == c.TDBG_OK) g_ctx = ctx;
}

/// (table generation, segment, column) blocks already resident in HBM.
/// Segments are immutable, so a staged block never needs invalidation; a
/// TRUNCATE or ALTER re-mints `cache_uid`, which is the GPU table uid.
const StagedKey = struct { uid: u64, segment_id: u64, column: u32 };
var g_staged_mu: std.Thread.Mutex = .{};
var g_staged: std.AutoHashMapUnmanaged(StagedKey, void) = .empty;

fn typeTag(t: Type) ?u32 {
    return switch (t) {
        .boolean => c.TDBG_BOOLEAN,
        .tinyint => c.TDBG_TINYINT,
        .smallint => c.TDBG_SMALLINT,
        .int => c.TDBG_INT,
        .bigint => c.TDBG_BIGINT,
        .date => c.TDBG_DATE,
        .datetime => c.TDBG_DATETIME,
        .decimal64 => c.TDBG_DECIMAL64,
        .float => c.TDBG_FLOAT,
        .double => c.TDBG_DOUBLE,
        else => null,
    };
}

fn isIntTag(tag: u32) bool {
    return tag != c.TDBG_FLOAT and tag != c.TDBG_DOUBLE;
}

fn mapStatus(rc: i32) anyerror {
    return switch (rc) {
        c.TDBG_ERR_OOM => error.MemoryBudgetExceeded, // src/memory.zig:31-35
        c.TDBG_ERR_CANCELLED => error.QueryCancelled, // src/memory.zig:60
        else => blk: {
            std.log.err("GPU group: {s}", .{std.mem.span(c.tdbg_last_error())});
            break :blk error.Unexpected;
        },
    };
}

fn columnIndex(table: *api.Table, name: []const u8) ?usize {
    for (table.schema.columns, 0..) |col, i| {
        if (types.columnNameEql(col.name, name)) return i;
    }
    return null;
}

// ---- the plan descriptor ---------------------------------------------------------

/// Arrays the tdbg_plan_desc points into; freed once tdbg_plan_create copied them.
const Desc = struct {
    keys: std.ArrayListUnmanaged(c.tdbg_key_part) = .empty,
    aggs: std.ArrayListUnmanaged(c.tdbg_agg_spec) = .empty,
    preds: std.ArrayListUnmanaged(c.tdbg_predicate) = .empty,
    where: std.ArrayListUnmanaged(c.tdbg_filter_node) = .empty,
    order: std.ArrayListUnmanaged(c.tdbg_order_spec) = .empty,
    having: std.ArrayListUnmanaged(c.tdbg_having_node) = .empty,
    columns: std.ArrayListUnmanaged(u32) = .empty, // table columns the plan reads (staged on demand)

    fn deinit(self: *Desc, a: Allocator) void {
        self.keys.deinit(a);
        self.aggs.deinit(a);
        self.preds.deinit(a);
        self.where.deinit(a);
        self.order.deinit(a);
        self.having.deinit(a);
        self.columns.deinit(a);
    }

    fn use(self: *Desc, a: Allocator, col: usize) !void {
        for (self.columns.items) |x| if (x == col) return;
        try self.columns.append(a, @intCast(col));
    }
};

/// A literal as the column's type holds it: i64 for the integer family
/// (a DECIMAL literal as its mantissa at the column's scale), f64 for floats.
const Lit = union(enum) { i: i64, f: f64 };

fn coerceLiteral(v: types.Value, col_type: Type) ?Lit {
    const tag = typeTag(col_type) orelse return null;
    if (!isIntTag(tag)) return switch (v) {
        .double => |x| .{ .f = x },
        .float => |x| .{ .f = x },
        .int => |x| .{ .f = @floatFromInt(x) },
        .bigint => |x| .{ .f = @floatFromInt(x) },
        else => null,
    };
    if (col_type == .decimal64) return switch (v) {
        .decimal64 => |x| .{ .i = x }, // the planner scaled it to the column
        else => null,
    };
    return switch (v) {
        .int => |x| .{ .i = x },
        .bigint => |x| .{ .i = x },
        .tinyint => |x| .{ .i = x },
        .smallint => |x| .{ .i = x },
        .date => |x| .{ .i = x },
        .datetime => |x| .{ .i = x },
        .boolean => |x| .{ .i = @intFromBool(x) },
        else => null,
    };
}

/// The WHERE: its top-level `col op literal` conjuncts become fused range
/// terms (`preds`); every other top-level conjunct is a postfix tree, and the
/// trees are AND-ed into one (`where`).  False: a form not covered.
fn whereTop(a: Allocator, table: *api.Table, d: *Desc, expr: PredicateExpr) !bool {
    var trees: u32 = 0;
    if (!try whereConjunct(a, table, d, expr, &trees)) return false;
    if (trees > 1) {
        var n = std.mem.zeroes(c.tdbg_filter_node);
        n.kind = c.TDBG_HAVING_AND;
        n.arity = trees;
        try d.where.append(a, n);
    }
    return d.where.items.len <= c.TDBG_MAX_WHERE;
}

fn whereConjunct(a: Allocator, table: *api.Table, d: *Desc, expr: PredicateExpr, trees: *u32) !bool {
    switch (expr) {
        .@"and" => |children| {
            for (children) |ch| if (!try whereConjunct(a, table, d, ch, trees)) return false;
            return true;
        },
        .leaf => |p| {
            if (p.as_boolean) return false;
            const ci = columnIndex(table, p.col) orelse return false;
            const col = table.schema.columns[ci];
            const lit = coerceLiteral(p.val, col.type) orelse return false;
            try d.use(a, ci);
            var pr = std.mem.zeroes(c.tdbg_predicate);
            pr.column_idx = @intCast(ci);
            pr.type = typeTag(col.type).?;
            pr.op = @intFromEnum(p.op);
            pr.nullable = @intFromBool(col.nullable);
            switch (lit) {
                .i => |x| pr.i64 = x,
                .f => |x| pr.f64 = x,
            }
            try d.preds.append(a, pr);
            return true;
        },
        else => {
            trees.* += 1;
            return wherePostfix(a, table, d, expr);
        },
    }
}

/// One WHERE subtree in postfix order (tdbg_filter_node).
fn wherePostfix(a: Allocator, table: *api.Table, d: *Desc, expr: PredicateExpr) !bool {
    var n = std.mem.zeroes(c.tdbg_filter_node);
    switch (expr) {
        .leaf => |p| {
            if (p.as_boolean) return false;
            const ci = columnIndex(table, p.col) orelse return false;
            const col = table.schema.columns[ci];
            const lit = coerceLiteral(p.val, col.type) orelse return false;
            try d.use(a, ci);
            n.kind = c.TDBG_HAVING_CMP;
            n.column_idx = @intCast(ci);
            n.type = typeTag(col.type).?;
            n.op = @intFromEnum(p.op);
            switch (lit) {
                .i => |x| n.i64 = x,
                .f => |x| n.f64 = x,
            }
        },
        .is_null, .is_not_null => |name| {
            const ci = columnIndex(table, name) orelse return false;
            try d.use(a, ci);
            n.kind = if (expr == .is_null) c.TDBG_HAVING_IS_NULL else c.TDBG_HAVING_IS_NOT_NULL;
            n.column_idx = @intCast(ci);
        },
        .not => |child| {
            if (!try wherePostfix(a, table, d, child.*)) return false;
            n.kind = c.TDBG_HAVING_NOT;
        },
        .@"and", .@"or" => |children| {
            for (children) |ch| if (!try wherePostfix(a, table, d, ch)) return false;
            n.kind = if (expr == .@"and") c.TDBG_HAVING_AND else c.TDBG_HAVING_OR;
            n.arity = @intCast(children.len);
        },
        else => return false, // LIKE, column-column, subquery and day leaves stay on the CPU
    }
    try d.where.append(a, n);
    return true;
}

/// HAVING in postfix order over the output columns; a literal in the output
/// column's domain (valueAsNum, v2_shape_group_topn.zig:2391-2407).
fn havingPart(a: Allocator, d: *Desc, expr: PredicateExpr, outputs: []const Column) !bool {
    switch (expr) {
        .leaf => |p| {
            const t = outputIndex(outputs, p.col) orelse return false;
            var h = std.mem.zeroes(c.tdbg_having_node);
            h.kind = c.TDBG_HAVING_CMP;
            h.target = @intCast(t);
            h.op = @intFromEnum(p.op);
            switch (p.val) {
                .double => |x| {
                    h.lit_float = 1;
                    h.lit_f64 = x;
                },
                .float => |x| {
                    h.lit_float = 1;
                    h.lit_f64 = x;
                },
                .int, .bigint, .tinyint, .smallint => {
                    const v: i128 = switch (p.val) {
                        .int => |x| x,
                        .bigint => |x| x,
                        .tinyint => |x| x,
                        .smallint => |x| x,
                        else => unreachable,
                    };
                    // a DECIMAL output compares mantissas at its scale
                    const scaled: i128 = if (outputs[t].type.decimalSpec()) |sp|
                        v * std.math.pow(i128, 10, sp.s)
                    else
                        v;
                    h.lit_lo = @bitCast(@as(u64, @truncate(@as(u128, @bitCast(scaled)))));
                    h.lit_hi = @truncate(scaled >> 64);
                },
                else => return false,
            }
            try d.having.append(a, h);
        },
        .is_null, .is_not_null => |name| {
            var h = std.mem.zeroes(c.tdbg_having_node);
            h.kind = if (expr == .is_null) c.TDBG_HAVING_IS_NULL else c.TDBG_HAVING_IS_NOT_NULL;
            h.target = @intCast(outputIndex(outputs, name) orelse return false);
            try d.having.append(a, h);
        },
        .not => |child| {
            if (!try havingPart(a, d, child.*, outputs)) return false;
            var h = std.mem.zeroes(c.tdbg_having_node);
            h.kind = c.TDBG_HAVING_NOT;
            try d.having.append(a, h);
        },
        .@"and", .@"or" => |children| {
            for (children) |ch| if (!try havingPart(a, d, ch, outputs)) return false;
            var h = std.mem.zeroes(c.tdbg_having_node);
            h.kind = if (expr == .@"and") c.TDBG_HAVING_AND else c.TDBG_HAVING_OR;
            h.arity = @intCast(children.len);
            try d.having.append(a, h);
        },
        else => return false,
    }
    return d.having.items.len <= c.TDBG_MAX_HAVING;
}

fn outputIndex(outputs: []const Column, name: []const u8) ?usize {
    for (outputs, 0..) |o, i| if (types.columnNameEql(o.name, name)) return i;
    return null;
}

// ---- the operator ---------------------------------------------------------------------

pub const GpuGroup = struct {
    allocator: Allocator,
    table: *api.Table,
    request: Request,
    plan: *c.tdbg_plan,
    columns: []u32, // table columns the plan reads
    col_types: []u32, // their tdbg types
    output_schema: []Column,
    output_cols: []ColumnStore, // owned copies, as LowCardGroup owns its output
    views: []ColumnView,
    est_groups: u64,
    row_count: usize = 0,
    emitted: bool = false,
    // The CPU core, built only when the snapshot next() captures holds rows
    // the GPU does not see (memtable rows, tombstones).
    fallback: ?exec.Query = null,

    pub fn next(self: *GpuGroup) !?exec.Batch {
        if (self.fallback) |*q| return q.next();
        if (self.emitted) return null;
        try exec.memory.checkCancelled(self.allocator);
        const table = self.table;
        table.ddl_lock.lockSharedUncancelable(table.io);
        defer table.ddl_lock.unlockShared(table.io);
        const snap = try Scan.captureSnapshotAlloc(table, self.allocator, exec.memory.accountantOf(self.allocator));
        defer snap.release(table);
        defer self.allocator.free(snap.segments);
        if (snap.memtable_row_count != 0 or snap.tomb_generation != 0) {
            // rows outside immutable segments: the CPU operator serves this snapshot
            self.fallback = (try v2_pipeline.tryBuildLowCardGroup(self.allocator, table, self.request)) orelse
                (try v2_pipeline.tryBuildGroupTopN(self.allocator, table, self.request)) orelse
                return error.Unexpected;
            return self.fallback.?.next();
        }
        // stage what HBM does not hold yet, then name every row group of the snapshot
        var rgs: std.ArrayListUnmanaged(c.tdbg_rg_ref) = .empty;
        defer rgs.deinit(self.allocator);
        for (snap.segments) |entry| {
            try self.stageSegment(entry.segment_id);
            for (0..entry.row_group_count) |rg| {
                try rgs.append(self.allocator, .{ .segment_id = entry.segment_id, .row_group_idx = @intCast(rg), .pad = 0 });
            }
        }
        try exec.memory.checkCancelled(self.allocator);
        var res: ?*c.tdbg_result = null;
        // through the process-wide device loop (gpu_serving.zig: continuous
        // batching, admission), else directly
        if (gpu_serving.enabled()) {
            res = try gpu_serving.execute(self.allocator, g_ctx.?, self.plan, rgs.items, mapStatus);
        }
        if (res == null) {
            const rc = c.tdbg_execute(self.plan, rgs.items.ptr, @intCast(rgs.items.len), &res);
            if (rc != c.TDBG_OK) return mapStatus(rc);
        }
        defer c.tdbg_result_free(res);
        try self.copyResult(res.?);
        self.emitted = true;
        for (self.output_cols, 0..) |col, i| self.views[i] = col.view();
        return .{ .schema = self.output_schema, .values = self.views, .row_count = self.row_count };
    }

    /// Reads the segment's .dat file once and stages the plan's columns of
    /// every row group (the decoded-block cache fill, segment_reader.zig:115).
    fn stageSegment(self: *GpuGroup, segment_id: u64) !void {
        const uid = self.table.cache_uid;
        {
            g_staged_mu.lock();
            defer g_staged_mu.unlock();
            var all = true;
            for (self.columns) |col| {
                if (!g_staged.contains(.{ .uid = uid, .segment_id = segment_id, .column = col })) all = false;
            }
            if (all) return;
        }
        var name_buf: [32]u8 = undefined;
        const name = try api.Table.segmentFileName(&name_buf, segment_id);
        const bytes = try self.table.segments_dir.readFileAlloc(self.table.io, name, self.allocator, .unlimited);
        defer self.allocator.free(bytes);
        var seg: ?*c.tdbg_segment = null;
        var rc = c.tdbg_segment_open(bytes.ptr, bytes.len, @intCast(self.table.schema.columns.len), &seg);
        if (rc != c.TDBG_OK) return mapStatus(rc);
        defer c.tdbg_segment_close(seg);
        var info: c.tdbg_segment_info = undefined;
        rc = c.tdbg_segment_get_info(seg, &info);
        if (rc != c.TDBG_OK) return mapStatus(rc);
        for (self.columns, self.col_types) |col, tag| {
            for (0..info.n_row_groups) |rg| {
                rc = c.tdbg_segment_stage(g_ctx, seg, uid, @intCast(rg), col, tag);
                if (rc != c.TDBG_OK) return mapStatus(rc);
            }
            g_staged_mu.lock();
            defer g_staged_mu.unlock();
            try g_staged.put(std.heap.page_allocator, .{ .uid = uid, .segment_id = segment_id, .column = col }, {});
        }
    }

    /// Copies the library's result columns into the operator's ColumnStores.
    fn copyResult(self: *GpuGroup, res: *c.tdbg_result) !void {
        const a = self.allocator;
        const n: usize = @intCast(c.tdbg_result_num_rows(res));
        self.row_count = n;
        for (self.output_cols, self.output_schema, 0..) |*col, schema_col, i| {
            var v: c.tdbg_column_view = undefined;
            const rc = c.tdbg_result_column(res, @intCast(i), &v);
            if (rc != c.TDBG_OK) return mapStatus(rc);
            const base: [*]const u8 = @ptrCast(v.data orelse @as(?*const anyopaque, @ptrCast(&v)));
            switch (schema_col.type) {
                .boolean => try col.data.boolean.appendSlice(a, base[0..n]),
                .tinyint => try col.data.tinyint.appendSlice(a, @as([*]const i8, @ptrCast(@alignCast(base)))[0..n]),
                .smallint => try col.data.smallint.appendSlice(a, @as([*]const i16, @ptrCast(@alignCast(base)))[0..n]),
                .int => try col.data.int.appendSlice(a, @as([*]const i32, @ptrCast(@alignCast(base)))[0..n]),
                .date => try col.data.date.appendSlice(a, @as([*]const i32, @ptrCast(@alignCast(base)))[0..n]),
                .bigint => try col.data.bigint.appendSlice(a, @as([*]const i64, @ptrCast(@alignCast(base)))[0..n]),
                .datetime => try col.data.datetime.appendSlice(a, @as([*]const i64, @ptrCast(@alignCast(base)))[0..n]),
                .decimal64 => try col.data.decimal64.appendSlice(a, @as([*]const i64, @ptrCast(@alignCast(base)))[0..n]),
                .decimal128 => try col.data.decimal128.appendSlice(a, @as([*]const i128, @ptrCast(@alignCast(base)))[0..n]),
                .float => try col.data.float.appendSlice(a, @as([*]const f32, @ptrCast(@alignCast(base)))[0..n]),
                .double => try col.data.double.appendSlice(a, @as([*]const f64, @ptrCast(@alignCast(base)))[0..n]),
                else => return error.TypeMismatch,
            }
            if (schema_col.nullable) {
                for (0..n) |r| {
                    const valid = if (v.validity) |bm| (bm[r >> 3] >> @intCast(r & 7)) & 1 != 0 else true;
                    try col.appendValidBit(a, r, valid);
                }
            }
        }
    }

    pub fn outputSchema(self: *GpuGroup) []const Column {
        if (self.fallback) |*q| return q.outputSchema();
        return self.output_schema;
    }

    pub fn addPrune(_: *GpuGroup, _: exec.Predicate) !void {} // zone maps prune inside tdbg_execute

    pub fn accountant(_: *GpuGroup) ?*exec.memory.MemoryAccountant {
        return null; // device memory is budgeted by the GPU context (TDBG_ERR_OOM)
    }

    pub fn stats(self: *GpuGroup) exec.PipelineStats {
        return .{ .upper_rows = self.est_groups };
    }

    /// The VTable's explain (exec.zig:409): one line naming the operator
    /// ("GPU group"), then the scan it reads, one level deeper.
    pub fn explain(self: *GpuGroup, out: *std.ArrayList(u8), allocator: Allocator, depth: usize) !void {
        try exec.explainIndent(out, allocator, depth);
        try out.appendSlice(allocator, "HashAggregate (GPU group: MI355X fused scan/filter/group-by/top-N)\n");
        try exec.explainIndent(out, allocator, depth + 1);
        try out.appendSlice(allocator, "Scan ");
        try out.appendSlice(allocator, self.table.name);
        try out.appendSlice(allocator, " (HBM-resident segments)\n");
    }

    /// Forwards a KILL to the device execution (cooperative: launched work finishes).
    pub fn cancel(self: *GpuGroup) void {
        _ = c.tdbg_cancel(self.plan);
    }

    pub fn deinit(self: *GpuGroup) void {
        const a = self.allocator;
        if (self.fallback) |*q| q.deinit();
        c.tdbg_plan_destroy(self.plan);
        for (self.output_cols) |*col| col.deinit(a);
        a.free(self.output_cols);
        a.free(self.views);
        a.free(self.output_schema);
        a.free(self.columns);
        a.free(self.col_types);
        a.destroy(self);
    }
};

/// Builds the GPU plan for the request; null = decline (the CPU core runs).
pub fn tryBuild(allocator: Allocator, table: *api.Table, request: Request) !?exec.Query {
    g_ctx_once.call();
    const ctx = g_ctx orelse return null; // no GPU
    // shapes this operator does not cover: derived keys or inputs, the WHERE's
    // conjuncts over computed columns (residual_filter, computed_where:
    // v2_shape_group_topn.zig:71-76), UDAFs
    if (request.derived.len != 0 or request.residual_filter != null or request.computed_where != null) return null;
    if (request.group_cols.len == 0 or request.group_cols.len > 8 or request.aggs.len == 0 or request.aggs.len > 16)
        return null;

    var d: Desc = .{};
    defer d.deinit(allocator);
    const n_out = request.group_cols.len + request.aggs.len;
    const outputs = try allocator.alloc(Column, n_out);
    var outputs_owned = true;
    defer if (outputs_owned) allocator.free(outputs);

    for (request.group_cols, 0..) |name, i| {
        const ci = columnIndex(table, name) orelse return null;
        const col = table.schema.columns[ci];
        const tag = typeTag(col.type) orelse return null;
        if (!isIntTag(tag)) return null; // float and string keys stay on the CPU
        try d.keys.append(allocator, .{ .column_idx = @intCast(ci), .type = tag });
        try d.use(allocator, ci);
        outputs[i] = .{ .name = col.name, .type = col.type, .nullable = col.nullable };
    }
    for (request.aggs, 0..) |agg, i| {
        var spec = std.mem.zeroes(c.tdbg_agg_spec);
        var in_type: ?Type = null;
        spec.op = switch (agg.func) {
            .count => if (agg.col == null) c.TDBG_AGG_COUNT_STAR else c.TDBG_AGG_COUNT,
            .count_distinct => c.TDBG_AGG_COUNT_DISTINCT,
            .sum => c.TDBG_AGG_SUM,
            .avg => c.TDBG_AGG_AVG,
            .min => c.TDBG_AGG_MIN,
            .max => c.TDBG_AGG_MAX,
            else => return null, // UDAFs and the other aggregates decline (engine_v2.zig:1002)
        };
        spec.column_idx = -1;
        if (agg.col) |name| {
            const ci = columnIndex(table, name) orelse return null;
            const col = table.schema.columns[ci];
            const tag = typeTag(col.type) orelse return null;
            spec.column_idx = @intCast(ci);
            spec.input_type = tag;
            spec.input_scale = if (col.type.decimalSpec()) |sp| sp.s else 0;
            spec.nullable = @intFromBool(col.nullable);
            try d.use(allocator, ci);
            in_type = col.type;
        }
        const out_type = aggregate.aggOutputTypeFor(agg, in_type) catch return null;
        const nullable = switch (agg.func) {
            .sum, .avg, .min, .max => true,
            else => false,
        };
        try d.aggs.append(allocator, spec);
        outputs[request.group_cols.len + i] = .{ .name = agg.as, .type = out_type, .nullable = nullable };
    }
    if (request.where_filter) |w| {
        if (!try whereTop(allocator, table, &d, w)) return null;
    }
    for (request.order_specs) |s| {
        const t = outputIndex(outputs, s.col) orelse return null;
        try d.order.append(allocator, .{ .target = @intCast(t), .desc = @intFromBool(s.desc) });
    }
    if (request.having_filter) |h| {
        if (!try havingPart(allocator, &d, h, outputs)) return null;
    }

    var desc = std.mem.zeroes(c.tdbg_plan_desc);
    desc.table_uid = table.cache_uid;
    desc.n_keys = @intCast(d.keys.items.len);
    desc.n_aggs = @intCast(d.aggs.items.len);
    desc.keys = d.keys.items.ptr;
    desc.aggs = d.aggs.items.ptr;
    desc.n_preds = @intCast(d.preds.items.len);
    desc.preds = d.preds.items.ptr;
    desc.n_order = @intCast(d.order.items.len);
    desc.order = d.order.items.ptr;
    desc.limit = request.limit;
    desc.offset = request.offset;
    desc.n_having = @intCast(d.having.items.len);
    desc.having = d.having.items.ptr;
    desc.n_where = @intCast(d.where.items.len);
    desc.where = d.where.items.ptr;
    var plan: ?*c.tdbg_plan = null;
    const rc = c.tdbg_plan_create(ctx, &desc, &plan);
    if (rc == c.TDBG_ERR_UNSUPPORTED) return null; // decline: the CPU core runs
    if (rc != c.TDBG_OK) return mapStatus(rc);
    errdefer c.tdbg_plan_destroy(plan);

    const op = try allocator.create(GpuGroup);
    errdefer allocator.destroy(op);
    const output_cols = try allocator.alloc(ColumnStore, n_out);
    var built: usize = 0;
    errdefer {
        for (output_cols[0..built]) |*col| col.deinit(allocator);
        allocator.free(output_cols);
    }
    for (outputs, 0..) |o, i| {
        output_cols[i] = try ColumnStore.initCapacity(allocator, o.type, o.nullable, 1, 0);
        built += 1;
    }
    const col_types = try allocator.alloc(u32, d.columns.items.len);
    for (d.columns.items, col_types) |ci, *t| t.* = typeTag(table.schema.columns[ci].type).?;
    op.* = .{
        .allocator = allocator,
        .table = table,
        .request = request,
        .plan = plan.?,
        .columns = try allocator.dupe(u32, d.columns.items),
        .col_types = col_types,
        .output_schema = outputs,
        .output_cols = output_cols,
        .views = try allocator.alloc(ColumnView, n_out),
        .est_groups = request.limit,
    };
    outputs_owned = false;
    return exec.makeQuery(allocator, op); // exec.zig:1221-1239
}

// ---- the seam ----------------------------------------------------------------------------
//
// src/exec/v2_pipeline.zig gains a sibling builder:
//
//     const gpu_group = @import("gpu_group.zig");
//     pub fn tryBuildGpuGroup(allocator: std.mem.Allocator, table: *api.Table,
//                             request: GroupTopNRequest) !?exec.Query {
//         return gpu_group.tryBuild(allocator, table, request);
//     }
//
// and engine_v2.buildGroupTopN (src/exec/engine_v2.zig:968) tries it first,
// where the low-cardinality core is tried today:
//
//     var built_q = try v2_pipeline.tryBuildGpuGroup(input.allocator, table, request);
//     if (built_q == null and residual == null and computed_where == null)
//         built_q = try v2_pipeline.tryBuildLowCardGroup(input.allocator, table, lowcard_request);
//     if (built_q != null) { ... } else built_q = try v2_pipeline.tryBuildGroupTopN(...);
//
// The GPU operator applies HAVING, ORDER BY and LIMIT itself (as the silo
// does), so `having_outside` stays false when it is the one built.
