"""Error paths through the C-ABI, as the reference tests them.

* A memory budget the query cannot fit: the execution fails with
  TDBG_ERR_OOM (error.MemoryBudgetExceeded) and nothing it took stays in use
  (tests/integration/v2_group_topn_test.zig:6-59: MemoryBudgetExceeded, then
  inUse() == 0); the context keeps working.
* Cancellation (memory.checkCancelled, src/memory.zig:60): a served request
  cancelled while queued returns TDBG_ERR_CANCELLED and never reaches the
  GPU, the others in the queue are unaffected; a cancel issued before an
  execution is asked for does not touch it; a cancel from another thread
  during an execution returns CANCELLED or a complete result, never a partial
  one, and the plan runs normally afterwards.  Every kernel exits normally:
  cancellation is checked on the host between launches.
"""
import threading

import numpy as np
import pytest

from helpers import check_with_oracle, rows_of
from thindb_amd import gen, gpu
from thindb_amd.gpu import Type
from thindb_amd.exec import AggSpec, Batch, Column, GpuSession, Request, SortSpec, Table, build_plan, output_schema

pytestmark = pytest.mark.gpu


def rows(table, req, cols):
    return rows_of(Batch(output_schema(table, req), cols, len(cols[0].values) if cols else 0))


def test_memory_budget_exceeded_releases_everything(oracle_lib):
    sess = GpuSession(0, hbm_budget_bytes=384 << 20)
    try:
        n = 1_200_000  # at least 2^24 / 16 rows: the 2^24-key span takes a dense table
        rng = np.random.default_rng(1)
        t = Table([Column("k", Type.BIGINT), Column("v", Type.BIGINT)],
                  {"k": rng.integers(0, 1 << 24, n), "v": rng.integers(-1000, 1000, n)})
        req = Request(["k"], [AggSpec("count", None, "c"), AggSpec("sum", "v", "s"), AggSpec("min", "v", "mn"),
                              AggSpec("max", "v", "mx")])
        plan, cols = build_plan(sess, t, req)
        sess.stage_table(t, cols)
        staged = sess.ctx.bytes_in_use()
        print(f"[errors] staged {staged >> 20} MiB of a 384 MiB budget", flush=True)
        # a 2^24-key dense table of 40-byte records (640 MiB) does not fit
        with pytest.raises(gpu.MemoryBudgetExceeded):
            plan.execute()
        plan.close()
        assert sess.ctx.bytes_in_use() == staged
        # the context still runs a query that fits, and gets it right
        small = Request(["k"], [AggSpec("count", None, "c")], [SortSpec("c", True), SortSpec("k")], 5)
        p2, _ = build_plan(sess, t, small)
        got = rows(t, small, p2.execute())
        p2.close()
        assert sess.ctx.bytes_in_use() == staged
        check_with_oracle(got, oracle_lib, t, small, 4)
    finally:
        sess.close()


def test_cancel_reaches_queued_requests(session, oracle_lib):
    big = gen.table("highcard", 20_000_000, seed=4)
    big_req = Request(["user_id", "bucket"], [AggSpec("count", None, "c"), AggSpec("sum", "v", "s")],
                      [SortSpec("c", True)], 10)
    small = [(gen.batched_table(31, i, 200_000), gen.batched_request(31, i)) for i in range(3)]
    plans = []
    for t, r in [(big, big_req)] + small:
        p, cols = build_plan(session, t, r)
        session.stage_table(t, cols)
        plans.append(p)
    srv = gpu.Server(session.ctx, lanes=1, max_batch=1)
    try:
        before = plans[1]
        before.cancel()  # asked before the request exists: it runs
        tickets = [srv.submit(p) for p in plans]
        plans[2].cancel()  # queued behind the multi-millisecond high-cardinality query
        print("[errors] cancelled a queued request", flush=True)
        outcomes = []
        for (t, r), tk in zip([(big, big_req)] + small, tickets):
            try:
                outcomes.append(rows(t, r, tk.wait()))
            except gpu.QueryCancelled:
                outcomes.append(None)
            tk.close()
        assert outcomes[2] is None, "the queued request was cancelled"
        for i in (0, 1, 3):
            assert outcomes[i] is not None
            t, r = ([(big, big_req)] + small)[i]
            if i:
                check_with_oracle(outcomes[i], oracle_lib, t, r, 8)
        # the plan runs normally afterwards
        t, r = small[1]
        check_with_oracle(rows(t, r, plans[2].execute()), oracle_lib, t, r, 8)
    finally:
        srv.stop()
        for p in plans:
            p.close()


def test_cancel_from_another_thread(session, oracle_lib):
    t = gen.table("highcard", 20_000_000, seed=5)
    req = Request(["user_id", "bucket"], [AggSpec("count", None, "c"), AggSpec("max", "w", "m")],
                  [SortSpec("c", True), SortSpec("m")], 10)
    plan, cols = build_plan(session, t, req)
    session.stage_table(t, cols)
    try:
        first = rows(t, req, plan.execute())  # warm: the kernels are loaded
        outcomes = []
        for delay in (0.0, 0.0005, 0.002):
            result = {}

            def run():
                try:
                    result["rows"] = rows(t, req, plan.execute())
                except gpu.QueryCancelled:
                    result["rows"] = None

            th = threading.Thread(target=run)
            th.start()
            threading.Event().wait(delay)
            plan.cancel()
            th.join(60)
            assert not th.is_alive()
            outcomes.append(result["rows"])
            if result["rows"] is not None:  # finished before the cancel: whole
                assert result["rows"] == first
        print(f"[errors] cross-thread cancels: {['cancelled' if o is None else 'finished' for o in outcomes]}",
              flush=True)
        # and the plan runs normally afterwards
        assert rows(t, req, plan.execute()) == first
    finally:
        plan.close()


def test_wait_in_slices_then_cancel(session, oracle_lib):
    """tdbg_ticket_wait_for: a connection thread waits in slices (PENDING while
    its request is queued behind a long one), checks for a KILL between
    slices and forwards it with tdbg_cancel; the request then ends cancelled
    or whole, never partial."""
    big = gen.table("highcard", 20_000_000, seed=6)
    big_req = Request(["user_id", "bucket"], [AggSpec("count", None, "c")], [SortSpec("c", True)], 10)
    t, r = gen.batched_table(32, 3, 200_000), gen.batched_request(32, 3)
    plans = []
    for tt, rr in ((big, big_req), (t, r)):
        p, cols = build_plan(session, tt, rr)
        session.stage_table(tt, cols)
        plans.append(p)
    srv = gpu.Server(session.ctx, lanes=1, max_batch=1)
    try:
        want = rows(t, r, plans[1].execute())
        check_with_oracle(want, oracle_lib, t, r, 8)
        for cancel in (False, True):
            tk_big = srv.submit(plans[0])
            tk = srv.submit(plans[1])
            slices, got = 0, None
            while True:
                try:
                    cols = tk.wait_for(200)
                except gpu.QueryCancelled:
                    got = "cancelled"
                    break
                if cols is not None:
                    got = rows(t, r, cols)
                    break
                slices += 1
                if cancel and slices == 1:
                    plans[1].cancel()  # the KILL seen between slices
            print(f"[errors] wait_for: {slices} pending slices, cancel={cancel} -> "
                  f"{got if got == 'cancelled' else 'whole result'}", flush=True)
            assert slices >= 1, "queued behind the long request"
            if not cancel:
                assert got == want
            else:
                assert got == "cancelled" or got == want
            tk_big.wait()
            tk_big.close()
            tk.close()
    finally:
        srv.stop()
        for p in plans:
            p.close()
