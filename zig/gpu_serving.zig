//! Wire-level serving for thindb-server: the process-wide GPU request loop
//! (tdbg_server in include/thindb_gpu.h) that connection threads submit their
//! grouped queries to.
//!
//! thindb-server runs every MySQL (and PG / native) connection on a thread of
//! its own (src/net/mysql/server.zig:191 spawns ConnJob.run per accepted
//! connection) and a grouped SELECT reaches the GPU operator (gpu_group.zig)
//! through engine_v2.buildGroupTopN on that thread.  With serving on,
//! GpuGroup.next() hands the execution to `execute` below instead of calling
//! tdbg_execute: the request joins the device loop's queue - small requests
//! that queue while the GPU is busy share one batched launch, large ones run
//! alone on their own lanes, admission keeps the estimated device bytes of
//! running work within a budget - and the connection thread waits on its
//! ticket.  Nothing changes in the accept loop: connection threads are the
//! submitters.
//!
//! A maintainer adds this file as src/exec/gpu_serving.zig beside
//! gpu_group.zig.  Zig is absent from this build's image, so it is written
//! against the reference's interfaces as they stand and not compiled here;
//! the same C-ABI calls (tdbg_server_submit_rgs, tdbg_ticket_wait_for,
//! tdbg_cancel) are exercised through ctypes by tests/test_serving.py and
//! tests/test_gpu_errors.py.  Reference interfaces used:
//!   - core_scheduler.global().parkAdmission() / Parked.unpark()
//!         src/util/core_scheduler.zig:107-117, :242-260 - a connection
//!         thread waiting on the GPU runs no CPU work, so its statement
//!         leaves the CPU fan-out sizing while it waits (its admission slot
//!         stays held), as a statement blocked on a slow client does
//!   - exec.memory.checkCancelled(allocator)   src/memory.zig:60 - KILL QUERY
//!         (handleQuery clears the flag per statement, src/net/mysql/
//!         server.zig:1102) is polled between bounded waits and forwarded to
//!         the device with tdbg_cancel
//!   - SET THINDB_...=  (trySetThindbEnvVar, src/net/mysql/server.zig:1085)
//!         sets the process environment, so THINDB_GPU_SERVE=0 turns serving
//!         off for later statements without a restart
//!   - the thindb-server shutdown path (src/cmd/server.zig: after the
//!         listener threads are joined) calls `shutdown`
const std = @import("std");
const Allocator = std.mem.Allocator;

const exec = @import("exec.zig");
const core_scheduler = @import("../util/core_scheduler.zig");

const c = @cImport(@cInclude("thindb_gpu.h"));

/// Microseconds per bounded wait: the KILL latency while a request is queued
/// or running (a cancel reaches queued requests before they launch, and a
/// running one between its kernels).
const wait_slice_us: u64 = 1000;

var g_server: ?*c.tdbg_server = null;
var g_mu: std.Thread.Mutex = .{};

fn envU64(name: [*:0]const u8, default: u64) u64 {
    const v = std.posix.getenv(std.mem.span(name)) orelse return default;
    return std.fmt.parseInt(u64, v, 10) catch default;
}

/// Serving is on unless THINDB_GPU_SERVE=0 (read per statement).
pub fn enabled() bool {
    const v = std.posix.getenv("THINDB_GPU_SERVE") orelse return true;
    return !std.mem.eql(u8, v, "0");
}

/// The process-wide device loop of `ctx`'s GPU, started on first use (as
/// core_scheduler.global() is created lazily).  Null when it cannot start;
/// the caller then executes directly.
fn server(ctx: *c.tdbg_context) ?*c.tdbg_server {
    g_mu.lock();
    defer g_mu.unlock();
    if (g_server) |s| return s;
    var cfg = std.mem.zeroes(c.tdbg_server_config);
    cfg.lanes = @intCast(envU64("THINDB_GPU_LANES", 0)); // 0: the library's default (3)
    cfg.max_batch = @intCast(envU64("THINDB_GPU_MAX_BATCH", 0));
    cfg.admit_bytes = envU64("THINDB_GPU_ADMIT_BYTES", 0); // 0: a quarter of the device budget
    cfg.large_rows = envU64("THINDB_GPU_LARGE_ROWS", 0);
    var s: ?*c.tdbg_server = null;
    if (c.tdbg_server_start(ctx, &cfg, &s) != c.TDBG_OK) {
        std.log.warn("GPU serving loop not started: {s}", .{std.mem.span(c.tdbg_last_error())});
        return null;
    }
    g_server = s;
    return s;
}

/// Stops the device loop after what is queued has run (thindb-server
/// shutdown, after the listener threads are joined).
pub fn shutdown() void {
    g_mu.lock();
    defer g_mu.unlock();
    if (g_server) |s| c.tdbg_server_stop(s);
    g_server = null;
}

/// Runs `plan` over the snapshot's row groups `rgs` through the device loop
/// and returns its result (the caller frees it with tdbg_result_free), or
/// null when the loop is unavailable (the caller calls tdbg_execute).
/// Errors map as tdbg_execute's do: error.QueryCancelled after a KILL,
/// error.MemoryBudgetExceeded, error.Unexpected.
pub fn execute(
    allocator: Allocator,
    ctx: *c.tdbg_context,
    plan: *c.tdbg_plan,
    rgs: []const c.tdbg_rg_ref,
    mapStatus: *const fn (i32) anyerror,
) !?*c.tdbg_result {
    const s = server(ctx) orelse return null;
    var ticket: ?*c.tdbg_ticket = null;
    const rc = c.tdbg_server_submit_rgs(s, plan, rgs.ptr, @intCast(rgs.len), &ticket);
    if (rc != c.TDBG_OK) return mapStatus(rc);
    defer c.tdbg_ticket_free(ticket);

    // waiting on the GPU is not CPU work: out of fan-out sizing meanwhile
    var parked = core_scheduler.global().parkAdmission();
    defer parked.unpark();

    var forwarded = false;
    while (true) {
        var res: ?*c.tdbg_result = null;
        const wrc = c.tdbg_ticket_wait_for(ticket, wait_slice_us, &res);
        if (wrc == c.TDBG_OK) return res;
        if (wrc != c.TDBG_PENDING) return mapStatus(wrc);
        // a KILL between slices: stop the device work, then keep waiting for
        // the ticket to settle (cancelled, or whole if it had finished)
        if (!forwarded) {
            exec.memory.checkCancelled(allocator) catch {
                _ = c.tdbg_cancel(plan);
                forwarded = true;
            };
        }
    }
}

// ---- where it is called ----------------------------------------------------------
//
// GpuGroup.next() (gpu_group.zig) runs its execution through `execute` when
// `enabled()`, falling back to tdbg_execute when the loop is unavailable; and
// src/cmd/server.zig, after the listener threads are joined at shutdown:
//
//     @import("../exec/gpu_serving.zig").shutdown();
