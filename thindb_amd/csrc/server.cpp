// server.cpp — the device request loop (tdbg_server_*): admission by
// estimated device bytes, continuous batching of small requests into the
// persistent batched launch, and lanes that keep several units of work on the
// GPU at once.
//
// Lane loop: take the queue head; if it is large (many rows, or ORDER BY /
// LIMIT / OFFSET) run it alone through the single-query operator, else take
// every small request queued behind it (up to max_batch, stopping at a large
// one) and run them as one batch.  With two lanes or more, lane 0 takes small
// requests only (the first ones queued, skipping large ones), so a run of
// large requests on the other lanes never holds small ones back.  A unit is
// admitted only while the estimated bytes of the running units stay within
// admit_bytes (a unit always runs when nothing else does, so an oversized
// request cannot starve).
//
// Hand-offs spin briefly before they sleep: an idle lane polls the queue and
// a waiting caller polls its ticket for a few tens of microseconds, about
// what a futex wake-up costs, so an unloaded server adds little to a
// request's latency.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>

#include <immintrin.h>

#include "runtime.h"

using namespace tdbg;
using namespace tdbg::rt;

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

constexpr int64_t kSpinNs = 40000;  // polling before a hand-off sleeps

// Polls `ready` for up to `ns` (kSpinNs by default); true when it held.
template <typename F>
bool spin_until(F ready, int64_t ns = kSpinNs) {
  const int64_t end = now_ns() + ns;
  do {
    if (ready()) return true;
    for (int i = 0; i < 32; i++) _mm_pause();
  } while (now_ns() < end);
  return ready();
}

}  // namespace

struct tdbg_ticket {
  tdbg_plan* plan = nullptr;
  bool all_rgs = true;            // scans every staged row group, else the ones in `rgs`
  std::vector<tdbg_rg_ref> rgs;
  uint64_t cancel_gen = 0;  // the plan's cancel generation at submission
  uint64_t est = 0;
  bool large = false;
  std::mutex m;
  std::condition_variable cv;
  bool done = false;
  std::atomic<bool> done_flag{false};  // `done`, readable without the lock (spin-waits)
  int32_t rc = TDBG_PENDING;
  std::string err;
  tdbg_result* result = nullptr;
  int64_t t_submit = 0, t_start = 0, t_done = 0;
  uint32_t batch_size = 0, lane = 0;
  std::atomic<int> refs{2};  // the caller's handle and the server's
};

struct tdbg_server {
  tdbg_context* ctx = nullptr;
  tdbg_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv;
  std::deque<tdbg_ticket*> queue;
  std::atomic<uint32_t> queued{0};  // queue.size(), readable without the lock (spin-waits)
  uint64_t admitted = 0;
  bool stopping = false;
  std::vector<std::thread> lanes;
  std::vector<tdbg_batch*> batches;  // one reusable batch per lane
  tdbg_server_counters stats{};
};

namespace {

void unref(tdbg_ticket* t) {
  if (t->refs.fetch_sub(1) == 1) {
    if (t->result) tdbg_result_free(t->result);
    delete t;
  }
}

void finish(tdbg_ticket* t, int32_t rc, tdbg_result* r, uint32_t batch_size, uint32_t lane) {
  {
    std::lock_guard<std::mutex> lk(t->m);
    t->rc = rc;
    if (rc != TDBG_OK) t->err = tdbg_last_error();
    t->result = r;
    t->batch_size = batch_size;
    t->lane = lane;
    t->t_done = now_ns();
    t->done = true;
    t->done_flag.store(true, std::memory_order_release);
  }
  t->cv.notify_all();
  unref(t);
}

// Runs one request alone through the single-query operator.
int32_t run_alone(tdbg_ticket* t, tdbg_result** r) {
  static const tdbg_rg_ref none{};  // an empty list scans nothing
  const tdbg_rg_ref* rgs = t->all_rgs ? nullptr : (t->rgs.empty() ? &none : t->rgs.data());
  return execute_plan(t->plan, rgs, (uint32_t)t->rgs.size(), true, r, t->cancel_gen);
}

// Estimated device bytes of one execution: the group table, radix partition
// rows when the key domain is wide and well filled, tiles and the result.
uint64_t estimate(tdbg_context* ctx, const tdbg_plan* pl, uint64_t* rows_out) {
  std::shared_lock<std::shared_mutex> lk(ctx->mu);
  *rows_out = 0;
  auto it = ctx->tables.find(pl->table_uid);
  if (it == ctx->tables.end()) return 0;
  const Table* t = it->second.get();
  uint64_t rows = 0;
  for (auto& rg : t->rgs) rows += rg.n_rows;
  *rows_out = rows;
  unsigned __int128 dom = 1;
  for (auto& k : pl->keys) {
    auto c = t->cols.find(k.column_idx);
    const uint64_t span = c != t->cols.end() && c->second.has_stats
                              ? (uint64_t)c->second.max_i - (uint64_t)c->second.min_i + 1
                              : ~0ull;
    dom *= span;
    if (dom > ((unsigned __int128)1 << 62)) dom = (unsigned __int128)1 << 62;
  }
  const uint64_t d = (uint64_t)dom;
  const uint64_t groups = std::max<uint64_t>(1, std::min<uint64_t>(pl->est_groups ? pl->est_groups : d, rows));
  const uint64_t rec = 16 + 24 * pl->aggs.size();
  const bool dense = d <= kDenseMax && (d <= kDenseFloor || d <= kDenseRowsX * std::max<uint64_t>(rows, 1));
  const uint64_t table = (dense ? d : next_pow2(groups + groups / 2 + 1)) * rec;
  const uint64_t radix = (d > 65536 && rows >= (1u << 20) && d / 4 <= rows) ? rows * 32 : 0;
  const uint64_t out = groups * 9 * (pl->keys.size() + pl->aggs.size());
  return table + radix + out + (rows / kTileRows + 1) * sizeof(Tile);
}

void lane_main(tdbg_server* s, uint32_t lane) {
  (void)hipSetDevice(s->ctx->device);
  std::vector<tdbg_ticket*> unit;
  std::vector<tdbg_plan*> plans;
  std::vector<const std::vector<tdbg_rg_ref>*> rgs;
  const bool small_only = lane == 0 && s->cfg.lanes >= 2;
  for (;;) {
    uint64_t est = 0;
    spin_until([&] { return s->queued.load(std::memory_order_relaxed) != 0; });
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        // the first request this lane may take: the head, or for the
        // small-request lane the first small one
        auto first = s->queue.begin();
        if (small_only)
          while (first != s->queue.end() && (*first)->large) ++first;
        if (first != s->queue.end()) {
          tdbg_ticket* head = *first;
          if (s->admitted != 0 && s->admitted + head->est > s->cfg.admit_bytes) {
            s->stats.admission_waits++;
            s->cv.wait(lk);  // a running unit frees memory
            continue;
          }
          unit.clear();
          if (head->large) {
            unit.push_back(head);
            est = head->est;
            s->queue.erase(first);
          } else {
            auto it = first;
            while (it != s->queue.end() && unit.size() < s->cfg.max_batch) {
              tdbg_ticket* t = *it;
              if (t->large) {
                if (small_only) { ++it; continue; }
                break;
              }
              if (!unit.empty() && s->admitted + est + t->est > s->cfg.admit_bytes) break;
              unit.push_back(t);
              est += t->est;
              it = s->queue.erase(it);
            }
          }
          s->queued.store((uint32_t)s->queue.size(), std::memory_order_relaxed);
          s->admitted += est;
          s->stats.peak_admitted = std::max(s->stats.peak_admitted, s->admitted);
          break;
        }
        if (s->stopping) return;
        s->cv.wait(lk);
      }
    }
    const int64_t t0 = now_ns();
    for (tdbg_ticket* t : unit) t->t_start = t0;
    // requests cancelled while queued end here, before any device work
    const size_t n_unit = unit.size();
    std::vector<tdbg_ticket*> live;
    for (tdbg_ticket* t : unit) {
      if (cancelled(t->plan, t->cancel_gen)) {
        fail(TDBG_ERR_CANCELLED, "cancelled while queued");
        finish(t, TDBG_ERR_CANCELLED, nullptr, 1, lane);
      } else {
        live.push_back(t);
      }
    }
    if (live.size() == 1) {
      tdbg_result* r = nullptr;
      const int32_t rc = run_alone(live[0], &r);
      finish(live[0], rc, rc == TDBG_OK ? r : nullptr, 1, lane);
    } else if (!live.empty()) {
      plans.clear();
      rgs.clear();
      for (tdbg_ticket* t : live) {
        plans.push_back(t->plan);
        rgs.push_back(&t->rgs);
      }
      int32_t rc = batch_run_plans(s->batches[lane], plans.data(), (uint32_t)plans.size(), rgs.data());
      for (size_t i = 0; i < live.size(); i++) {
        tdbg_result* r = nullptr;
        int32_t ri = rc;
        if (rc == TDBG_OK) ri = batch_take_result(s->batches[lane], (uint32_t)i, &r);
        // a plan a batch cannot hold (a key domain only hash partitions
        // cover): each request runs alone
        if (rc == TDBG_ERR_UNSUPPORTED) ri = run_alone(live[i], &r);
        if (ri == TDBG_OK && cancelled(live[i]->plan, live[i]->cancel_gen)) {  // cancelled while it ran
          tdbg_result_free(r);
          r = nullptr;
          ri = fail(TDBG_ERR_CANCELLED, "cancelled");
        }
        finish(live[i], ri, ri == TDBG_OK ? r : nullptr, (uint32_t)live.size(), lane);
      }
    }
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->admitted -= est;
      s->stats.completed += n_unit;
      s->stats.units++;
      if (live.size() > 1) s->stats.batched += live.size();
    }
    s->cv.notify_all();
  }
}

}  // namespace

extern "C" {

int32_t tdbg_server_start(tdbg_context* ctx, const tdbg_server_config* cfg, tdbg_server** out) {
  if (!ctx || !out) return fail(TDBG_ERR_INVALID, "null argument");
  *out = nullptr;
  std::unique_ptr<tdbg_server> s(new tdbg_server());
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  if (!s->cfg.lanes) s->cfg.lanes = 3;
  if (!s->cfg.max_batch) s->cfg.max_batch = 64;
  if (!s->cfg.admit_bytes) s->cfg.admit_bytes = ctx->budget / 4;
  if (!s->cfg.large_rows) s->cfg.large_rows = 4u << 20;
  if (s->cfg.lanes > 16) return fail(TDBG_ERR_INVALID, "%u lanes (at most 16)", s->cfg.lanes);
  for (uint32_t i = 0; i < s->cfg.lanes; i++) {
    tdbg_batch* b = nullptr;
    int32_t rc = batch_open(ctx, &b);
    if (rc) {
      for (tdbg_batch* x : s->batches) batch_close(x);
      return rc;
    }
    s->batches.push_back(b);
  }
  for (uint32_t i = 0; i < s->cfg.lanes; i++) s->lanes.emplace_back(lane_main, s.get(), i);
  *out = s.release();
  return TDBG_OK;
}

int32_t tdbg_server_submit(tdbg_server* s, tdbg_plan* plan, tdbg_ticket** out) {
  return tdbg_server_submit_rgs(s, plan, nullptr, 0, out);
}

int32_t tdbg_server_submit_rgs(tdbg_server* s, tdbg_plan* plan, const tdbg_rg_ref* rgs, uint32_t n_rgs,
                               tdbg_ticket** out) {
  if (!s || !plan || !out || (n_rgs && !rgs)) return fail(TDBG_ERR_INVALID, "null argument");
  *out = nullptr;
  if (plan->ctx != s->ctx) return fail(TDBG_ERR_INVALID, "plan of another context");
  auto* t = new tdbg_ticket();
  t->plan = plan;
  t->all_rgs = rgs == nullptr;
  if (rgs) t->rgs.assign(rgs, rgs + n_rgs);
  t->cancel_gen = plan->cancel_gen.load();  // a tdbg_cancel from here on reaches this request
  t->t_submit = now_ns();
  uint64_t rows = 0;
  t->est = estimate(s->ctx, plan, &rows);
  // (an empty row-group list runs alone: a batch reads an empty list as every row group)
  t->large = rows >= s->cfg.large_rows || !plan_batchable(plan) || (!t->all_rgs && t->rgs.empty());
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) {
      delete t;
      return fail(TDBG_ERR_INVALID, "server is stopping");
    }
    s->queue.push_back(t);
    s->queued.store((uint32_t)s->queue.size(), std::memory_order_relaxed);
    s->stats.submitted++;
  }
  // every lane: one woken alone could be the small-request lane, which
  // leaves a large request to the others
  s->cv.notify_all();
  *out = t;
  return TDBG_OK;
}

int32_t tdbg_ticket_poll(tdbg_ticket* t) {
  if (!t) return fail(TDBG_ERR_INVALID, "null ticket");
  std::lock_guard<std::mutex> lk(t->m);
  if (!t->done) return TDBG_PENDING;
  if (t->rc != TDBG_OK) fail(t->rc, "%s", t->err.c_str());
  return t->rc;
}

int32_t tdbg_ticket_wait(tdbg_ticket* t, tdbg_result** out) {
  if (!t || !out) return fail(TDBG_ERR_INVALID, "null argument");
  *out = nullptr;
  spin_until([&] { return t->done_flag.load(std::memory_order_acquire); });
  std::unique_lock<std::mutex> lk(t->m);
  t->cv.wait(lk, [&] { return t->done; });
  if (t->rc != TDBG_OK) return fail(t->rc, "%s", t->err.c_str());
  *out = t->result;
  t->result = nullptr;
  return TDBG_OK;
}

int32_t tdbg_ticket_wait_for(tdbg_ticket* t, uint64_t timeout_us, tdbg_result** out) {
  if (!t || !out) return fail(TDBG_ERR_INVALID, "null argument");
  *out = nullptr;
  const int64_t budget = (int64_t)std::min<uint64_t>(timeout_us, 1ull << 40) * 1000;
  spin_until([&] { return t->done_flag.load(std::memory_order_acquire); }, std::min<int64_t>(budget, kSpinNs));
  std::unique_lock<std::mutex> lk(t->m);
  if (!t->cv.wait_for(lk, std::chrono::nanoseconds(std::max<int64_t>(0, budget - kSpinNs)), [&] { return t->done; }))
    return TDBG_PENDING;
  if (t->rc != TDBG_OK) return fail(t->rc, "%s", t->err.c_str());
  *out = t->result;
  t->result = nullptr;
  return TDBG_OK;
}

int32_t tdbg_ticket_info(const tdbg_ticket* t, tdbg_ticket_report* out) {
  if (!t || !out) return fail(TDBG_ERR_INVALID, "null argument");
  auto* mt = const_cast<tdbg_ticket*>(t);
  std::lock_guard<std::mutex> lk(mt->m);
  *out = tdbg_ticket_report{};
  out->submit_ns = t->t_submit;
  out->start_ns = t->t_start;
  out->done_ns = t->t_done;
  out->est_bytes = t->est;
  out->batch_size = t->batch_size;
  out->lane = t->lane;
  out->status = t->done ? t->rc : TDBG_PENDING;
  return TDBG_OK;
}

void tdbg_ticket_free(tdbg_ticket* t) {
  if (t) unref(t);
}

int32_t tdbg_server_stats(const tdbg_server* s, tdbg_server_counters* out) {
  if (!s || !out) return fail(TDBG_ERR_INVALID, "null argument");
  auto* ms = const_cast<tdbg_server*>(s);
  std::lock_guard<std::mutex> lk(ms->mu);
  *out = s->stats;
  return TDBG_OK;
}

void tdbg_server_stop(tdbg_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv.notify_all();
  for (auto& th : s->lanes) th.join();
  for (tdbg_batch* b : s->batches) batch_close(b);
  delete s;
}

}  // extern "C"
