// batch.hip's host side: many plans become one set of work items over one
// device layout, launched as a short fixed sequence (counters, record init,
// the persistent kernel) that is captured once as a HIP graph and replayed.
// The persistent kernel writes every plan's result rows straight into pinned
// host memory as the plan finishes; only the header (group counts, error
// flags) is copied afterwards.
//
// The autotuner times the candidate strategies (LDS direct table, HBM
// atomics, LDS hash table, radix partitions at two leaf sizes) on the batch's
// own plans, per bucket of (rows, estimated groups) in powers of four, and
// keeps the fastest per bucket in the context.
#include <algorithm>
#include <cstring>

#include "runtime.h"

using namespace tdbg;
using namespace tdbg::rt;

namespace {

constexpr uint64_t kItemRows = 1ull << 17;     // rows per scan work item
constexpr uint64_t kRadixItemRows = 1ull << 18;  // rows per radix work item (THINDB_GPU_BATCH_RADIX_ROWS)
constexpr uint32_t kSpinLimit = 1u << 21;      // polls (~seconds) before a dependency wait gives up
constexpr uint32_t kBatchLds = 78u * 1024u;    // LDS per workgroup: two workgroups per CU
constexpr uint32_t kTuneSample = 64;           // plans per bucket the autotuner times together
constexpr uint64_t kTuneRows = 128ull << 20;   // rows a tuning trial runs (the sample repeated)
constexpr uint32_t kTuneCopies = 16;           // repeats of the sample, at most
constexpr uint32_t kRowsPerSlot = 256;         // scan rows per LDS slot flushed, at least, per item
constexpr uint32_t kWaves = 1;                 // waves of a large batch (THINDB_GPU_BATCH_WAVES: more)
constexpr uint32_t kWavePlansMin = 32;         // plans per wave, at least
constexpr size_t kJitStrips = 8;               // plan structures a run-time compiled batch kernel specialises

uint32_t ilog2(uint64_t x) { return x ? 63u - (uint32_t)__builtin_clzll(x) : 0u; }
uint32_t bucket_of(uint64_t rows, uint64_t groups) { return (ilog2(rows) / 2) << 8 | (ilog2(groups) / 2); }

struct DevBuf {
  void* p = nullptr;
  uint64_t bytes = 0;
};

// Estimated groups of a plan: the hint, else the key domain, at most the rows.
uint64_t estimate_groups(const tdbg_plan* pl, Table* t, uint64_t rows) {
  if (pl->est_groups) return std::max<uint64_t>(1, std::min<uint64_t>(pl->est_groups, rows));
  unsigned __int128 dom = 1;
  for (auto& k : pl->keys) {
    auto it = t->cols.find(k.column_idx);
    if (it == t->cols.end() || !it->second.has_stats) return std::max<uint64_t>(1, rows);
    dom *= (unsigned __int128)((uint64_t)it->second.max_i - (uint64_t)it->second.min_i + 1);
    if (dom > rows) return std::max<uint64_t>(1, rows);
  }
  return std::max<uint64_t>(1, (uint64_t)dom);
}

PlanHint batch_hint(PlanHint h) {
  h.lds_budget = kBatchLds;
  h.max_bits1 = 10;
  h.single_level = true;
  h.dense_out = true;
  const char* env = getenv("THINDB_GPU_BATCH_WIDE_LEAVES");
  h.wide_leaves = !(env && env[0] == '0');
  return h;
}

}  // namespace

struct tdbg_batch {
  tdbg_context* ctx = nullptr;
  std::vector<tdbg_plan*> plans;
  // per plan, the row groups it scans (empty: every staged row group; a
  // request of the serving loop names its snapshot's)
  std::vector<std::vector<tdbg_rg_ref>> plan_rgs;
  std::vector<PlanHint> forced;  // autotuner trial: this hint for every plan
  bool use_graph = true;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::mutex mu;
  // built state
  bool built = false;
  std::vector<std::pair<uint64_t, uint64_t>> versions;  // table uid and version the build saw
  std::vector<tdbg_batch_route> info;
  std::vector<std::vector<uint64_t>> data_off, valid_off;
  uint64_t rows = 0, bytes = 0;
  uint32_t n_items = 0, n_static = 0, n_cblocks = 0, grid = 0, lds_bytes = 0, emit_cap = 0;
  uint64_t zero_words = 0, out_bytes = 0, err_off = 0, head_bytes = 0, result_bytes = 0;
  uint64_t sched_off = 0, slots_off[kBatchQueues] = {0, 0, 0};
  uint32_t q_items[kBatchQueues] = {0, 0, 0};
  DevBuf d_q, d_items, d_tiles, d_rec, d_rows, d_misc, d_zero, d_cblocks, d_head, d_run;
  PinnedBuf h_out;                 // results keep the buffer they view
  uint8_t* h_run = nullptr;        // pinned: the output base the kernel writes through
  uint8_t* h_desc = nullptr;       // pinned staging of the descriptors
  uint64_t h_desc_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  tdbg_batch_stats last{};
  // waves: a large batch runs as sub-batches over contiguous plan ranges on
  // the same stream; each wave's result copy, on copy_stream, overlaps the
  // kernels of the waves after it
  std::vector<tdbg_batch*> waves;
  std::vector<uint32_t> wave_first;  // first plan of each wave, then the plan count
  hipStream_t copy_stream = nullptr;
  hipEvent_t kernels_done = nullptr;
  hipEvent_t ev_kernels = nullptr;   // timing: the kernels ended (before the result copy)
  hipFunction_t jit = nullptr;       // the batch kernel compiled for these plans (nullptr: ahead of time)
  float kernels_ms = 0;              // last execution: device time before the result copy
};

namespace {

// Grows `d` to at least `bytes` (buffers are reused across builds).
int32_t dev_take(tdbg_batch* b, DevBuf& d, uint64_t bytes) {
  if (d.p && d.bytes >= bytes) return TDBG_OK;
  dev_free(b->ctx, d.p, d.bytes);
  d = DevBuf{};
  const uint64_t want = std::max<uint64_t>(bytes, 256);
  int32_t rc = dev_alloc(b->ctx, &d.p, want);
  if (rc) return rc;
  d.bytes = want;
  return TDBG_OK;
}

// Drops the captured graph (the plan must be rebuilt); buffers stay.
void release_built(tdbg_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->built = false;
}

void release_buffers(tdbg_batch* b) {
  release_built(b);
  for (DevBuf* d : {&b->d_q, &b->d_items, &b->d_tiles, &b->d_rec, &b->d_rows, &b->d_misc, &b->d_zero, &b->d_cblocks,
                    &b->d_head, &b->d_run}) {
    dev_free(b->ctx, d->p, d->bytes);
    *d = DevBuf{};
  }
  b->built = false;
}

template <typename T>
T* at(const DevBuf& d, uint64_t off) {
  return reinterpret_cast<T*>(static_cast<uint8_t*>(d.p) + off);
}

// The kernel sequence of one execution (captured into the graph).
int32_t enqueue(tdbg_batch* b, bool timed) {
  hipStream_t s = b->stream;
  HIP_TRY(launch_fill_u32(static_cast<uint32_t*>(b->d_zero.p), b->zero_words, 0, s), "batch counters");
  HIP_TRY(launch_fill_u32(at<uint32_t>(b->d_head, b->err_off), 16, 0, s), "batch error flags");
  HIP_TRY(launch_batch_init(static_cast<const BatchQuery*>(b->d_q.p), static_cast<const BatchBlock*>(b->d_cblocks.p),
                            b->n_cblocks, s), "batch record init");
  if (timed) HIP_TRY(hipEventRecord(b->ev[1], s), "event");
  if (b->jit && b->n_items)
    HIP_TRY(launch_batch_jit(b->jit, static_cast<const BatchQuery*>(b->d_q.p),
                             static_cast<const BatchItem*>(b->d_items.p), static_cast<const BatchArgs*>(b->d_run.p),
                             b->grid, b->lds_bytes, s), "batch kernel (compiled for its plans)");
  else
    HIP_TRY(launch_batch(static_cast<const BatchQuery*>(b->d_q.p), static_cast<const BatchItem*>(b->d_items.p),
                         static_cast<const BatchArgs*>(b->d_run.p), b->n_items, b->grid, b->lds_bytes, s),
            "batch kernel");
  if (timed) HIP_TRY(hipEventRecord(b->ev[2], s), "event");
  return TDBG_OK;
}

// Plans every member, lays out the device buffers and the work queue, and
// captures the graph.  Caller holds b->mu.
int32_t build(tdbg_batch* b) {
  release_built(b);
  tdbg_context* ctx = b->ctx;
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  std::shared_lock<std::shared_mutex> lk(ctx->mu);
  const uint32_t n = (uint32_t)b->plans.size();
  std::vector<BatchQuery> Q(n);
  std::vector<ScanSet> sets(n);
  std::vector<uint64_t> gmax(n);  // output rows upper bound per plan
  b->versions.clear();
  b->info.assign(n, tdbg_batch_route{});
  b->rows = b->bytes = 0;
  uint32_t lds_need = 4096;
  for (uint32_t i = 0; i < n; i++) {
    tdbg_plan* pl = b->plans[i];
    auto tit = ctx->tables.find(pl->table_uid);
    if (tit == ctx->tables.end())
      return fail(TDBG_ERR_NOT_STAGED, "plan %u: table %llu has no staged columns", i, (unsigned long long)pl->table_uid);
    Table* t = tit->second.get();
    b->versions.emplace_back(pl->table_uid, t->version);
    ScanSet& S = sets[i];
    const bool named = i < b->plan_rgs.size() && !b->plan_rgs[i].empty();
    int32_t rc = collect_scan(pl, t, named ? b->plan_rgs[i].data() : nullptr,
                              named ? (uint32_t)b->plan_rgs[i].size() : 0u, S);
    if (rc) return rc;
    PlanHint h;
    // THINDB_GPU_BATCH_STRATEGY=direct|global|hash|radix forces one strategy
    // for every plan (tests cover each path with it)
    const char* senv = getenv("THINDB_GPU_BATCH_STRATEGY");
    if (!b->forced.empty()) {
      h = b->forced[0];
    } else if (senv && senv[0]) {
      const std::string v(senv);
      h.strategy = v == "direct" ? S_DIRECT : v == "global" ? S_GLOBAL : v == "hash" ? S_LDS_HASH
                 : v == "radix" ? S_RADIX : S_AUTO;
    } else {
      std::lock_guard<std::mutex> tk(ctx->tune_mu);
      auto it = ctx->tuned.find(bucket_of(S.rows, estimate_groups(pl, t, S.rows)));
      if (it != ctx->tuned.end()) h = it->second;
    }
    h = batch_hint(h);
    if (S.rows == 0) h.strategy = S_GLOBAL;
    std::lock_guard<std::mutex> plk(pl->mu);
    if ((rc = prepare(pl, t, S, h))) return rc;
    if (pl->prep.radix && (!pl->prep.rx.dense_out || S.rows >= (1ull << 32))) {
      // a sparse key domain (compacted leaf output) or a table beyond 32-bit
      // row offsets: not partitioned in a batch
      h.strategy = S_GLOBAL;
      if ((rc = prepare(pl, t, S, h))) return rc;
    }
    const Prepared& P = pl->prep;
    if (P.tmpl.size() > (size_t)kTmplWords) return fail(TDBG_ERR_UNSUPPORTED, "plan %u: record wider than a batch allows", i);
    BatchQuery& q = Q[i];
    memset(&q, 0, sizeof q);
    q.sp = P.sp;
    q.rx = P.rx;
    q.rx.having = nullptr;  // batched leaves write dense records; HAVING joins the compaction
    q.op = P.op;
    q.fp = P.fp;
    q.n_having = (uint32_t)pl->having.size();
    q.having = static_cast<const HavingNode*>(pl->having_nodes.p);
    for (size_t w = 0; w < P.tmpl.size(); w++) q.tmpl[w] = P.tmpl[w];
    q.init_records = P.radix ? 0u : 1u;
    if (P.radix) q.rx.tile_rows = std::min<uint32_t>(q.rx.tile_rows, kBatchThreads * kVec);
    if (q.sp.g_cap >= (1ull << 32)) return fail(TDBG_ERR_UNSUPPORTED, "plan %u: more group records than a batch indexes", i);
    gmax[i] = std::max<uint64_t>(1, std::min<uint64_t>(q.sp.g_cap, S.rows));
    tdbg_batch_route& inf = b->info[i];
    inf.strategy = P.radix ? 4u : (q.sp.lds_mode == LDS_DIRECT ? 1u : (q.sp.lds_mode == LDS_HASH ? 3u : 2u));
    inf.lds_parts = q.sp.lds_mode == LDS_DIRECT ? q.sp.lds_parts : 1u;
    inf.radix_bits = P.radix ? q.rx.bits1 : 0u;
    if (q.sp.lds_mode != LDS_NONE) lds_need = std::max(lds_need, q.sp.lds_bytes);
    if (P.radix) lds_need = std::max<uint32_t>(lds_need, kRxStageOff + q.rx.tile_rows * q.rx.row.stride);
    b->rows += S.rows;
    b->bytes += S.bytes;
  }
  if (lds_need > kBatchLds) return fail(TDBG_ERR_DEVICE, "batch LDS layout of %u bytes", lds_need);
  // emit items stage a slab of records (kBatchThreads of them) in LDS when it fits
  for (uint32_t i = 0; i < n; i++) {
    const uint64_t stage = kEmitStageOff + (uint64_t)kBatchThreads * Q[i].sp.g_rec;
    if (stage <= kBatchLds) lds_need = std::max<uint32_t>(lds_need, (uint32_t)stage);
  }
  b->lds_bytes = lds_need;
  // device layout
  const char* rx_env = getenv("THINDB_GPU_BATCH_RADIX_ROWS");
  const uint64_t rx_rows = rx_env && rx_env[0] ? std::max<uint64_t>(kTileRows, strtoull(rx_env, nullptr, 10)) : kRadixItemRows;
  const uint32_t rx_chunk = (uint32_t)std::max<uint64_t>(1, rx_rows / kTileRows);
  uint64_t rec_b = 0, tile_b = 0, rows_b = 0, misc_b = 0;
  std::vector<uint64_t> rec_o(n), tile_o(n), rows_o(n), misc_o(n);
  for (uint32_t i = 0; i < n; i++) {
    const BatchQuery& q = Q[i];
    rec_o[i] = rec_b;
    rec_b += align_up(q.sp.g_cap * q.sp.g_rec, 256);
    tile_o[i] = tile_b;
    tile_b += align_up(sets[i].tiles.size() * sizeof(Tile), 256);
    if (b->info[i].strategy == 4u) {
      rows_o[i] = rows_b;
      rows_b += align_up(sets[i].rows * q.rx.row.stride + 256, 256);
      // bucket starts and ends, then per-item digit counts (one item per
      // kItemRows rows of tiles)
      const uint64_t n_hist = (sets[i].tiles.size() + rx_chunk - 1) / rx_chunk;
      misc_o[i] = misc_b;
      misc_b += align_up(((2ull << q.rx.bits1) + 2) * 8, 256) + align_up(n_hist * (4ull << q.rx.bits1), 256);
    }
  }
  // output: group counts and error flags (device header, copied to the
  // host after the kernel), then each plan's columns (host memory, written
  // by the kernel's emit items)
  b->err_off = align_up(8ull * n, 64);
  b->head_bytes = b->err_off + 64;
  uint64_t out_b = align_up(b->head_bytes, 256);
  b->data_off.assign(n, {});
  b->valid_off.assign(n, {});
  b->result_bytes = 0;
  for (uint32_t i = 0; i < n; i++) {
    FinalParams& fp = Q[i].fp;
    out_b = align_up(out_b, 256);
    b->data_off[i].assign(fp.n_outs, 0);
    b->valid_off[i].assign(fp.n_outs, ~0ull);
    for (uint32_t c = 0; c < fp.n_outs; c++) {
      OutDesc& o = fp.outs[c];
      out_b = align_up(out_b, 16);
      o.data_off = out_b;
      b->data_off[i][c] = out_b;
      out_b += gmax[i] * o.bytes;
      if (o.valid_off != ~0ull) {
        out_b = align_up(out_b, 16);
        o.valid_off = out_b;
        b->valid_off[i][c] = out_b;
        out_b += gmax[i];
      }
    }
  }
  b->out_bytes = out_b;
  // scan plans of one structure share a strip generated for it in a batch
  // kernel compiled at run time, for batches that are executed again and
  // again (tdbg_batch_create): not for the autotuner's trials, whose
  // strategies change from trial to trial, nor for the serving loop's
  // batches, whose plan sets change from launch to launch (a compile takes
  // about two seconds).  The structures are sorted, so one set compiles once.
  b->jit = nullptr;
  for (uint32_t i = 0; i < n; i++) Q[i].jit_sig = ~0u;
  if (b->use_graph && b->forced.empty()) {
    std::vector<std::pair<std::string, uint32_t>> keyed;  // (structure, plan)
    bool generic = false;
    for (uint32_t i = 0; i < n; i++) {
      if (b->info[i].strategy == 4u) continue;
      if (Q[i].sp.n_cols > 8) generic = true;
      else keyed.emplace_back(jit_signature(Q[i].sp), i);
    }
    std::sort(keyed.begin(), keyed.end());
    std::vector<ScanParams> sigs;
    for (size_t j = 0; j < keyed.size(); j++) {
      if (j == 0 || keyed[j].first != keyed[j - 1].first) {
        if (sigs.size() == kJitStrips) {
          generic = true;
          break;
        }
        sigs.push_back(Q[keyed[j].second].sp);
      }
      Q[keyed[j].second].jit_sig = (uint32_t)sigs.size() - 1;
    }
    if (!sigs.empty()) b->jit = batch_jit_kernel(sigs, generic);
    if (!b->jit)
      for (uint32_t i = 0; i < n; i++) Q[i].jit_sig = ~0u;
  }
  // record-init blocks and work items: histograms and scans are ready at
  // launch (the static queue); each plan's scatters, leaves and emit items
  // follow it in the item array, appended to the ready queues on the device
  // when the stage before completes
  std::vector<BatchBlock> cblocks;
  std::vector<BatchItem> hist_items, scan_items, dyn_items;
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; i++) order[i] = i;
  // radix plans with the most records first: their rows take longest to
  // reach the host, so their emit items should start early
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return Q[x].sp.g_cap > Q[y].sp.g_cap; });
  const uint32_t base_chunk = (uint32_t)std::max<uint64_t>(1, kItemRows / kTileRows);
  for (uint32_t i : order) {
    BatchQuery& q = Q[i];
    if (q.init_records)
      for (uint64_t r = 0; r < q.sp.g_cap; r += kBatchChunk) cblocks.push_back(BatchBlock{i, 0, r});
    const uint32_t n_tiles = (uint32_t)sets[i].tiles.size();
    uint32_t n_items_plan = 0;
    q.n_hist = q.n_scatter = q.n_comp = 0;
    if (b->info[i].strategy == 4u) {
      q.sc0 = (uint32_t)dyn_items.size();
      for (uint32_t t0 = 0, k = 0; t0 < n_tiles; t0 += rx_chunk, k++) {
        const uint32_t t1 = std::min(n_tiles, t0 + rx_chunk);
        hist_items.push_back(BatchItem{BI_RX_HIST, i, t0, t1, k, 0});  // part: the item's index
        dyn_items.push_back(BatchItem{BI_RX_SCATTER, i, t0, t1, k, 0});
        q.n_hist++;
        q.n_scatter++;
      }
      q.ag0 = (uint32_t)dyn_items.size();
      const uint64_t per = std::max<uint64_t>(1, (uint64_t)q.rx.n_leaves * rx_rows / std::max<uint64_t>(sets[i].rows, 1));
      for (uint64_t l0 = 0; l0 < q.rx.n_leaves; l0 += per) {
        dyn_items.push_back(BatchItem{BI_RX_AGG, i, (uint32_t)l0, (uint32_t)std::min<uint64_t>(q.rx.n_leaves, l0 + per), 0, 0});
        q.n_comp++;
      }
      if (q.n_hist == 0 || q.n_comp == 0)
        return fail(TDBG_ERR_DEVICE, "plan %u: radix partitioning without tiles or leaves", i);
      n_items_plan = q.n_hist + q.n_scatter + q.n_comp;
    } else {
      const uint32_t parts = q.sp.lds_mode == LDS_DIRECT ? q.sp.lds_parts : 1u;
      // an LDS table item ends with a flush of every slot to HBM (about five
      // device atomics each): give it at least kRowsPerSlot rows per slot so
      // the flush stays a small part of the item
      uint32_t scan_chunk = base_chunk;
      if (q.sp.lds_mode != LDS_NONE)
        scan_chunk = (uint32_t)std::max<uint64_t>(scan_chunk,
                                                  ((uint64_t)kRowsPerSlot * q.sp.lds_cap + kTileRows - 1) / kTileRows);
      for (uint32_t t0 = 0; t0 < n_tiles; t0 += scan_chunk)
        for (uint32_t p = 0; p < parts; p++) {
          scan_items.push_back(BatchItem{BI_SCAN, i, t0, std::min(n_tiles, t0 + scan_chunk), p, 0});
          q.n_comp++;
        }
      if (q.n_comp == 0) {  // nothing to scan: one empty item releases the emit items
        scan_items.push_back(BatchItem{BI_SCAN, i, 0, 0, 0, 0});
        q.n_comp = 1;
      }
      n_items_plan = q.n_comp;
    }
    q.em0 = (uint32_t)dyn_items.size();
    q.n_emit = 0;
    for (uint64_t r = 0; r < q.sp.g_cap; r += kEmitChunk, q.n_emit++)
      dyn_items.push_back(BatchItem{BI_EMIT, i, (uint32_t)r, (uint32_t)std::min<uint64_t>(q.sp.g_cap, r + kEmitChunk),
                                    q.n_emit, 0});
    if (q.n_emit == 0) return fail(TDBG_ERR_DEVICE, "plan %u: no group records", i);
    b->info[i].items = n_items_plan + q.n_emit;
  }
  // the largest scans first, so the tail of the static queue is short items
  std::stable_sort(scan_items.begin(), scan_items.end(), [](const BatchItem& x, const BatchItem& y) {
    return (x.a1 - x.a0) > (y.a1 - y.a0);
  });
  std::vector<BatchItem> items;
  items.reserve(hist_items.size() + scan_items.size() + dyn_items.size());
  items.insert(items.end(), hist_items.begin(), hist_items.end());
  items.insert(items.end(), scan_items.begin(), scan_items.end());
  b->n_static = (uint32_t)items.size();
  items.insert(items.end(), dyn_items.begin(), dyn_items.end());
  uint32_t per_queue[kBatchQueues] = {0, 0, 0};
  uint64_t n_emit_all = 0;
  for (uint32_t i = 0; i < n; i++) {
    Q[i].sc0 += b->n_static;
    Q[i].ag0 += b->n_static;
    Q[i].em0 += b->n_static;
    per_queue[BQ_SCATTER] += Q[i].n_scatter;
    per_queue[BQ_AGG] += b->info[i].strategy == 4u ? Q[i].n_comp : 0u;
    per_queue[BQ_EMIT] += Q[i].n_emit;
    n_emit_all += Q[i].n_emit;
  }
  for (uint32_t k = 0; k < kBatchQueues; k++) b->q_items[k] = per_queue[k];
  b->n_items = (uint32_t)items.size();
  b->n_cblocks = (uint32_t)cblocks.size();
  const uint32_t n_cu = (uint32_t)std::max(ctx->n_cu, 8);
  b->grid = std::max<uint32_t>(1, std::min<uint32_t>(n_cu * 2, b->n_items));
  const char* ecap = getenv("THINDB_GPU_BATCH_EMITTERS");
  b->emit_cap = ecap && ecap[0] ? (uint32_t)std::max(1, atoi(ecap)) : std::max<uint32_t>(1, b->grid * 3 / 32);
  // zeroed per execution: queue state, per-plan stage counters, ready-queue
  // slots, emit item counts
  uint64_t zero_b = 0;
  b->sched_off = zero_b;
  zero_b += align_up(sizeof(BatchSched), 64);
  const uint64_t sync_off = zero_b;
  zero_b += align_up(16ull * n, 64);
  for (uint32_t k = 0; k < kBatchQueues; k++) {
    b->slots_off[k] = zero_b;
    zero_b += align_up(4ull * per_queue[k], 64);
  }
  const uint64_t status_off = zero_b;
  zero_b += align_up(8ull * n_emit_all, 64);
  zero_b = align_up(zero_b, 256);
  // allocations
  int32_t rc;
  if ((rc = dev_take(b, b->d_q, sizeof(BatchQuery) * n)) || (rc = dev_take(b, b->d_items, sizeof(BatchItem) * items.size())) ||
      (rc = dev_take(b, b->d_tiles, tile_b)) || (rc = dev_take(b, b->d_rec, rec_b)) || (rc = dev_take(b, b->d_rows, rows_b)) ||
      (rc = dev_take(b, b->d_misc, misc_b)) || (rc = dev_take(b, b->d_zero, zero_b)) ||
      (rc = dev_take(b, b->d_cblocks, sizeof(BatchBlock) * std::max<size_t>(1, cblocks.size()))) ||
      (rc = dev_take(b, b->d_head, b->head_bytes)) || (rc = dev_take(b, b->d_run, 256))) {
    release_buffers(b);
    return rc;
  }
  b->zero_words = zero_b / 4;
  // device pointers into the layout
  uint64_t status_at = status_off;
  for (uint32_t i = 0; i < n; i++) {
    BatchQuery& q = Q[i];
    q.sp.g_state = at<uint8_t>(b->d_rec, rec_o[i]);
    q.sp.tiles = at<const Tile>(b->d_tiles, tile_o[i]);
    q.sp.n_tiles = (uint32_t)sets[i].tiles.size();
    q.sp.err = at<uint32_t>(b->d_head, b->err_off);
    q.op.g_state = q.sp.g_state;
    q.fp.g_state = q.sp.g_state;
    q.sync = at<uint32_t>(b->d_zero, sync_off + 16ull * i);
    q.emit_status = at<uint64_t>(b->d_zero, status_at);
    status_at += 8ull * q.n_emit;
    q.n_out = at<uint64_t>(b->d_head, 8ull * i);
    if (b->info[i].strategy == 4u) {
      const uint64_t nb = 1ull << q.rx.bits1;
      q.item_hist = at<uint32_t>(b->d_misc, misc_o[i] + align_up(((2ull << q.rx.bits1) + 2) * 8, 256));
      q.rx.rows1 = at<uint8_t>(b->d_rows, rows_o[i]);
      q.rx.rows2 = q.rx.rows1;
      q.rx.digits1 = nullptr;  // one level in a batch
      q.rx.bucket_start = at<uint64_t>(b->d_misc, misc_o[i]);
      q.rx.bucket_fill = q.rx.bucket_start + nb + 1;
      q.rx.leaf_start = nullptr;
    }
  }
  // descriptors: staged back to back in pinned memory, copied asynchronously
  // on the batch's stream ahead of the kernels that read them
  const uint64_t q_b = sizeof(BatchQuery) * n, it_b = sizeof(BatchItem) * items.size();
  const uint64_t cb_b = sizeof(BatchBlock) * cblocks.size();
  const uint64_t desc_b = align_up(tile_b, 256) + align_up(q_b, 256) + align_up(it_b, 256) + cb_b;
  if (b->h_desc_cap < desc_b) {
    if (b->h_desc) (void)hipHostFree(b->h_desc);
    b->h_desc = nullptr;
    b->h_desc_cap = 0;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&b->h_desc), desc_b, hipHostMallocDefault), "pinned descriptors");
    b->h_desc_cap = desc_b;
  }
  uint8_t* h = b->h_desc;
  for (uint32_t i = 0; i < n; i++)
    if (!sets[i].tiles.empty()) memcpy(h + tile_o[i], sets[i].tiles.data(), sets[i].tiles.size() * sizeof(Tile));
  uint64_t o = align_up(tile_b, 256);
  auto put = [&](DevBuf& d, const void* src, uint64_t bytes, uint64_t at_off) -> hipError_t {
    if (!bytes) return hipSuccess;
    if (src) memcpy(h + at_off, src, bytes);
    return hipMemcpyAsync(d.p, h + at_off, bytes, hipMemcpyHostToDevice, b->stream);
  };
  HIP_TRY(put(b->d_tiles, nullptr, tile_b, 0), "batch tiles");
  HIP_TRY(put(b->d_q, Q.data(), q_b, o), "batch plans");
  o += align_up(q_b, 256);
  HIP_TRY(put(b->d_items, items.data(), it_b, o), "batch items");
  o += align_up(it_b, 256);
  HIP_TRY(put(b->d_cblocks, cblocks.data(), cb_b, o), "batch blocks");
  b->built = true;
  // kernel attributes are set outside any capture
  HIP_TRY(launch_batch(nullptr, nullptr, nullptr, 0, 1, 0, b->stream), "batch kernel setup");
  const char* genv = getenv("THINDB_GPU_GRAPH");
  if (b->use_graph && !(genv && genv[0] == '0')) {
    HIP_TRY(hipStreamSynchronize(b->stream), "descriptors");  // the capture starts on an idle stream
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "graph capture");
    rc = enqueue(b, false);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIP_TRY(ce, "graph capture");
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
  }
  return TDBG_OK;
}

// A member table was restaged, evicted or dropped since the build.  Caller
// holds ctx->mu.
bool stale(const tdbg_batch* b) {
  for (auto& v : b->versions) {
    auto it = b->ctx->tables.find(v.first);
    if (it == b->ctx->tables.end() || it->second->version != v.second) return true;
  }
  return false;
}

// Builds the batch if it is stale or new and makes sure its pinned result
// buffer is free.  Caller holds b->mu (or that of the batch it is a wave of).
int32_t prepare_run(tdbg_batch* b) {
  int32_t rc;
  {
    std::shared_lock<std::shared_mutex> lk(b->ctx->mu);
    if (b->built && stale(b)) b->built = false;
  }
  if (!b->built && (rc = build(b))) return rc;
  if (!b->h_out.p || b->h_out.p.use_count() > 1 || b->h_out.cap < b->out_bytes) {
    // results of an earlier execution still view the old buffer: keep it
    b->h_out = PinnedBuf{};
    if ((rc = pinned_acquire(b->ctx->pinned, b->out_bytes, b->h_out))) return rc;
  }
  if (!b->h_run) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&b->h_run), 256, hipHostMallocDefault), "pinned run block");
  return TDBG_OK;
}

// Enqueues the kernels on b->stream and the result copy on `cs` (after the
// kernels).  Caller holds ctx->mu shared until the copy completed.
int32_t launch(tdbg_batch* b, hipStream_t cs) {
  int32_t rc;
  hipStream_t s = b->stream;
  if (stale(b)) return fail(TDBG_ERR_NOT_STAGED, "batch: a member table changed while the batch was planned");
  // the kernel's arguments, with the device's address of the pinned output
  // buffer it writes result rows through, go to d_run before every launch
  // (the graph keeps only d_run's address)
  void* dev_out = nullptr;
  HIP_TRY(hipHostGetDevicePointer(&dev_out, b->h_out.p.get(), 0), "pinned output address");
  BatchArgs a{};
  a.Q = static_cast<const BatchQuery*>(b->d_q.p);
  a.items = static_cast<const BatchItem*>(b->d_items.p);
  a.n_static = b->n_static;
  for (uint32_t k = 0; k < kBatchQueues; k++) a.q_items[k] = b->q_items[k];
  a.sched = at<BatchSched>(b->d_zero, b->sched_off);
  for (uint32_t k = 0; k < kBatchQueues; k++) a.slots[k] = at<uint32_t>(b->d_zero, b->slots_off[k]);
  a.err = at<uint32_t>(b->d_head, b->err_off);
  a.out = static_cast<uint8_t*>(dev_out);
  a.spin_limit = kSpinLimit;
  a.emit_cap = b->emit_cap;
  a.lds_bytes = b->lds_bytes;
  static_assert(sizeof(BatchArgs) <= 256, "d_run holds the kernel arguments");
  memcpy(b->h_run, &a, sizeof a);
  HIP_TRY(hipEventRecord(b->ev[0], s), "event");
  HIP_TRY(hipMemcpyAsync(b->d_run.p, b->h_run, sizeof a, hipMemcpyHostToDevice, s), "batch arguments");
  if (b->exec) HIP_TRY(hipGraphLaunch(b->exec, s), "graph launch");
  else if ((rc = enqueue(b, true))) return rc;
  HIP_TRY(hipEventRecord(b->ev_kernels, s), "event");
  if (cs != s) {
    HIP_TRY(hipEventRecord(b->kernels_done, s), "event");
    HIP_TRY(hipStreamWaitEvent(cs, b->kernels_done, 0), "copy stream wait");
  }
  HIP_TRY(hipMemcpyAsync(b->h_out.p.get(), b->d_head.p, b->head_bytes, hipMemcpyDeviceToHost, cs), "result header");
  HIP_TRY(hipEventRecord(b->ev[3], cs), "event");
  return TDBG_OK;
}

// After the copy: the error flags, then the statistics.
int32_t finish(tdbg_batch* b) {
  const uint32_t* err = reinterpret_cast<const uint32_t*>(b->h_out.p.get() + b->err_off);
  uint32_t bits = 0;
  for (int i = 0; i < 16; i++) bits |= err[i];
  if (bits & ERR_BATCH_WAIT) return fail(TDBG_ERR_DEVICE, "batch: a dependency wait gave up");
  if (bits & ERR_KEY_RANGE) return fail(TDBG_ERR_INVALID, "batch: a key value lies outside the staged zone-map statistics");
  if (bits & ERR_HASH_FULL) return fail(TDBG_ERR_OOM, "batch: a group table is full (est_groups too small)");
  const bool graph = b->exec != nullptr;
  float total = 0, scan = 0;
  hipEventElapsedTime(&total, b->ev[0], b->ev[3]);
  hipEventElapsedTime(&b->kernels_ms, b->ev[0], b->ev_kernels);
  if (!graph) hipEventElapsedTime(&scan, b->ev[1], b->ev[2]);
  b->last = tdbg_batch_stats{};
  b->last.total_ms = total;
  b->last.scan_ms = scan;
  b->last.rows = b->rows;
  b->last.bytes = b->bytes;
  b->last.result_bytes = b->out_bytes;
  b->last.n_plans = (uint32_t)b->plans.size();
  b->last.n_items = b->n_items;
  b->last.graph = graph ? 1u : 0u;
  return TDBG_OK;
}

// One execution into the pinned result buffer(s).  Caller holds b->mu.
int32_t run(tdbg_batch* b) {
  int32_t rc;
  if (b->waves.empty()) {
    if ((rc = prepare_run(b))) return rc;
    HIP_TRY(hipSetDevice(b->ctx->device), "hipSetDevice");
    std::shared_lock<std::shared_mutex> lk(b->ctx->mu);  // staged columns stay put while the kernels read them
    if ((rc = launch(b, b->stream))) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream), "batch");
    return finish(b);
  }
  for (tdbg_batch* w : b->waves)
    if ((rc = prepare_run(w))) return rc;
  HIP_TRY(hipSetDevice(b->ctx->device), "hipSetDevice");
  {
    std::shared_lock<std::shared_mutex> lk(b->ctx->mu);
    HIP_TRY(hipEventRecord(b->ev[0], b->stream), "event");
    for (tdbg_batch* w : b->waves)
      if ((rc = launch(w, b->copy_stream))) {
        (void)hipStreamSynchronize(b->copy_stream);
        (void)hipStreamSynchronize(b->stream);
        return rc;
      }
    HIP_TRY(hipEventRecord(b->ev[3], b->copy_stream), "event");
    HIP_TRY(hipStreamSynchronize(b->copy_stream), "batch copies");
    HIP_TRY(hipStreamSynchronize(b->stream), "batch");
  }
  b->last = tdbg_batch_stats{};
  b->last.graph = 1u;
  b->last.waves = (uint32_t)b->waves.size();
  for (tdbg_batch* w : b->waves) {
    if ((rc = finish(w))) return rc;
    b->last.scan_ms += w->last.scan_ms;
    b->last.rows += w->last.rows;
    b->last.bytes += w->last.bytes;
    b->last.result_bytes += w->last.result_bytes;
    b->last.n_plans += w->last.n_plans;
    b->last.n_items += w->last.n_items;
    b->last.graph &= w->last.graph;
  }
  float total = 0;
  hipEventElapsedTime(&total, b->ev[0], b->ev[3]);
  b->last.total_ms = total;
  b->built = true;
  return TDBG_OK;
}

// The wave holding plan i (and i's index in it), or b itself.
tdbg_batch* wave_of(tdbg_batch* b, uint32_t& i) {
  if (b->waves.empty()) return b;
  const auto it = std::upper_bound(b->wave_first.begin(), b->wave_first.end(), i);
  const size_t w = (size_t)(it - b->wave_first.begin()) - 1;
  i -= b->wave_first[w];
  return b->waves[w];
}

int32_t make_batch(tdbg_context* ctx, tdbg_plan* const* plans, uint32_t n, hipStream_t stream, tdbg_batch** out) {
  std::unique_ptr<tdbg_batch> b(new tdbg_batch());
  b->ctx = ctx;
  b->plans.assign(plans, plans + n);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  if (stream) {
    b->stream = stream;
  } else {
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
    b->own_stream = true;
  }
  for (auto& e : b->ev) HIP_TRY(hipEventCreate(&e), "event");
  HIP_TRY(hipEventCreateWithFlags(&b->kernels_done, hipEventDisableTiming), "event");
  HIP_TRY(hipEventCreate(&b->ev_kernels), "event");
  *out = b.release();
  return TDBG_OK;
}

void free_batch(tdbg_batch* b);

// Splits a large batch into waves of about equal staged rows (contiguous
// plan ranges).  THINDB_GPU_BATCH_WAVES sets their number (1: no waves).
int32_t split_waves(tdbg_batch* b) {
  const char* env = getenv("THINDB_GPU_BATCH_WAVES");
  const uint32_t n = (uint32_t)b->plans.size();
  uint32_t w = env && env[0] ? (uint32_t)std::max(1, atoi(env)) : kWaves;
  w = std::min<uint32_t>(w, n / kWavePlansMin);
  if (w <= 1) return TDBG_OK;
  std::vector<uint64_t> rows(n, 0);
  uint64_t total = 0;
  {
    std::shared_lock<std::shared_mutex> lk(b->ctx->mu);
    for (uint32_t i = 0; i < n; i++) {
      auto it = b->ctx->tables.find(b->plans[i]->table_uid);
      rows[i] = it == b->ctx->tables.end() ? 0 : it->second->next_row;
      total += rows[i];
    }
  }
  HIP_TRY(hipStreamCreateWithFlags(&b->copy_stream, hipStreamNonBlocking), "copy stream");
  uint64_t acc = 0;
  uint32_t first = 0;
  for (uint32_t i = 0; i < n; i++) {
    acc += rows[i];
    const bool cut = (uint64_t)(b->waves.size() + 1) * total <= acc * w || i == n - 1;
    if (cut && b->waves.size() + 1 < w && i + 1 < n || i == n - 1) {
      tdbg_batch* sub = nullptr;
      int32_t rc = make_batch(b->ctx, b->plans.data() + first, i + 1 - first, b->stream, &sub);
      if (rc) return rc;
      sub->use_graph = b->use_graph;
      b->waves.push_back(sub);
      b->wave_first.push_back(first);
      first = i + 1;
    }
  }
  b->wave_first.push_back(n);
  return TDBG_OK;
}

void free_batch(tdbg_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->copy_stream) (void)hipStreamSynchronize(b->copy_stream);
  for (tdbg_batch* w : b->waves) free_batch(w);
  if (b->copy_stream) (void)hipStreamDestroy(b->copy_stream);
  if (b->kernels_done) (void)hipEventDestroy(b->kernels_done);
  if (b->ev_kernels) (void)hipEventDestroy(b->ev_kernels);
  release_buffers(b);
  if (b->h_desc) (void)hipHostFree(b->h_desc);
  if (b->h_run) (void)hipHostFree(b->h_run);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->own_stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // namespace

namespace tdbg {
namespace rt {

int32_t batch_open(tdbg_context* ctx, tdbg_batch** out) {
  tdbg_plan* none = nullptr;
  int32_t rc = make_batch(ctx, &none, 0, nullptr, out);
  if (rc == TDBG_OK) (*out)->use_graph = false;
  return rc;
}

int32_t batch_run_plans(tdbg_batch* b, tdbg_plan* const* plans, uint32_t n,
                        const std::vector<tdbg_rg_ref>* const* rgs) {
  std::lock_guard<std::mutex> lk(b->mu);
  b->plans.assign(plans, plans + n);
  b->plan_rgs.assign(n, {});
  if (rgs)
    for (uint32_t i = 0; i < n; i++)
      if (rgs[i]) b->plan_rgs[i] = *rgs[i];
  release_built(b);
  return run(b);
}

int32_t batch_take_result(tdbg_batch* b, uint32_t i, tdbg_result** out) {
  std::lock_guard<std::mutex> lk(b->mu);
  if (!b->built || !b->h_out.p || i >= b->plans.size()) return fail(TDBG_ERR_INVALID, "no result %u", i);
  const uint8_t* h = b->h_out.p.get();
  const uint64_t n_out = reinterpret_cast<const uint64_t*>(h)[i];
  return make_result(b->plans[i], n_out, h, b->data_off[i], b->valid_off[i],
                     std::shared_ptr<void>(b->h_out.p, b->h_out.p.get()), out);
}

void batch_close(tdbg_batch* b) { free_batch(b); }

}  // namespace rt
}  // namespace tdbg

extern "C" {

int32_t tdbg_batch_create(tdbg_context* ctx, tdbg_plan* const* plans, uint32_t n_plans, tdbg_batch** out) {
  if (!ctx || !out || (n_plans && !plans)) return fail(TDBG_ERR_INVALID, "null argument");
  *out = nullptr;
  if (n_plans == 0) return fail(TDBG_ERR_INVALID, "empty batch");
  for (uint32_t i = 0; i < n_plans; i++) {
    const tdbg_plan* pl = plans[i];
    if (!pl || pl->ctx != ctx) return fail(TDBG_ERR_INVALID, "plan %u is null or of another context", i);
    if (!plan_batchable(pl))
      return fail(TDBG_ERR_UNSUPPORTED, "plan %u has ORDER BY / LIMIT / OFFSET or COUNT(DISTINCT): run it alone", i);
  }
  int32_t rc = make_batch(ctx, plans, n_plans, nullptr, out);
  if (rc == TDBG_OK && (rc = split_waves(*out))) {
    free_batch(*out);
    *out = nullptr;
  }
  return rc;
}

int32_t tdbg_batch_execute(tdbg_batch* b) {
  if (!b) return fail(TDBG_ERR_INVALID, "null batch");
  std::lock_guard<std::mutex> lk(b->mu);
  return run(b);
}

int32_t tdbg_batch_autotune(tdbg_batch* b, uint32_t reps) {
  if (!b) return fail(TDBG_ERR_INVALID, "null batch");
  std::lock_guard<std::mutex> lk(b->mu);
  tdbg_context* ctx = b->ctx;
  reps = std::max<uint32_t>(1, reps);
  // the batch's buckets
  std::map<uint32_t, std::vector<tdbg_plan*>> buckets;
  std::map<const tdbg_plan*, uint64_t> plan_rows;
  {
    std::shared_lock<std::shared_mutex> clk(ctx->mu);
    for (tdbg_plan* pl : b->plans) {
      auto tit = ctx->tables.find(pl->table_uid);
      if (tit == ctx->tables.end()) return fail(TDBG_ERR_NOT_STAGED, "table %llu has no staged columns",
                                                (unsigned long long)pl->table_uid);
      ScanSet S;
      int32_t rc = collect_scan(pl, tit->second.get(), nullptr, 0, S);
      if (rc) return rc;
      buckets[bucket_of(S.rows, estimate_groups(pl, tit->second.get(), S.rows))].push_back(pl);
      plan_rows[pl] = S.rows;
    }
  }
  const PlanHint cands[] = {
      PlanHint{S_DIRECT}, PlanHint{S_GLOBAL}, PlanHint{S_LDS_HASH}, PlanHint{S_RADIX, 0}, PlanHint{S_RADIX, 10},
  };
  for (auto& kv : buckets) {
    {
      std::lock_guard<std::mutex> tk(ctx->tune_mu);
      if (ctx->tuned.count(kv.first)) continue;
    }
    std::vector<tdbg_plan*> sample(kv.second.begin(),
                                   kv.second.begin() + std::min<size_t>(kv.second.size(), kTuneSample));
    // the sample repeated until it fills the GPU (about kTuneRows rows): in
    // a batch the plans of a bucket run among many others, so throughput
    // under load decides, not one small sample's latency
    uint64_t sample_rows = 0;
    for (tdbg_plan* pl : sample) sample_rows += plan_rows[pl];
    const size_t copies = (size_t)std::min<uint64_t>(kTuneCopies, std::max<uint64_t>(1, kTuneRows / std::max<uint64_t>(sample_rows, 1)));
    const size_t n_sample = sample.size();
    for (size_t c = 1; c < copies; c++) sample.insert(sample.end(), sample.begin(), sample.begin() + n_sample);
    double best = 1e30;
    PlanHint pick{};
    // the bucket's fewest groups: 4^(its group class)
    const uint64_t groups_lo = 1ull << (2 * (kv.first & 255u));
    for (const PlanHint& c : cands) {
      // An LDS hash table of a batch holds about kBatchLds / 40 groups (a
      // key, a count and three 8-byte states per slot); the planner takes it
      // for at most 8x that many.  Beyond, groups overflow the table all the
      // time: a trial whose time is mostly its result rows' PCIe stores can
      // still pick it, but among other plans it is several times slower than
      // radix partitions.
      if (c.strategy == S_LDS_HASH && groups_lo > 8ull * (kBatchLds / 40)) continue;
      tdbg_batch* sub = nullptr;
      if (make_batch(ctx, sample.data(), (uint32_t)sample.size(), b->stream, &sub)) continue;
      sub->forced = {c};
      sub->use_graph = false;
      double t = 1e30;
      bool ok = true;
      for (uint32_t r = 0; r <= reps && ok; r++) {
        if (run(sub)) {
          ok = false;
          break;
        }
        // the first run warms up; the result copy is the same for every
        // strategy, so the device time before it decides
        if (r > 0) t = std::min<double>(t, sub->kernels_ms);
      }
      free_batch(sub);
      if (ok && t < best) {
        best = t;
        pick = c;
      }
    }
    std::lock_guard<std::mutex> tk(ctx->tune_mu);
    ctx->tuned[kv.first] = pick;
  }
  release_built(b);  // re-plans with the tuned strategies at the next execution
  for (tdbg_batch* w : b->waves) release_built(w);
  return TDBG_OK;
}

int32_t tdbg_batch_result(tdbg_batch* b, uint32_t i, tdbg_result** out) {
  if (!b || !out) return fail(TDBG_ERR_INVALID, "null argument");
  *out = nullptr;
  std::lock_guard<std::mutex> lk(b->mu);
  if (i >= b->plans.size()) return fail(TDBG_ERR_INVALID, "no result %u", i);
  tdbg_batch* w = wave_of(b, i);
  if (!w->built || !w->h_out.p) return fail(TDBG_ERR_INVALID, "no result %u", i);
  const uint64_t n_out = reinterpret_cast<const uint64_t*>(w->h_out.p.get())[i];
  std::shared_ptr<void> keep(w->h_out.p, w->h_out.p.get());
  return make_result(w->plans[i], n_out, w->h_out.p.get(), w->data_off[i], w->valid_off[i], keep, out);
}

int32_t tdbg_batch_last_stats(const tdbg_batch* b, tdbg_batch_stats* out) {
  if (!b || !out) return fail(TDBG_ERR_INVALID, "null argument");
  *out = b->last;
  return TDBG_OK;
}

int32_t tdbg_batch_plan_info(const tdbg_batch* b, uint32_t i, tdbg_batch_route* out) {
  if (!b || !out) return fail(TDBG_ERR_INVALID, "null argument");
  if (i >= b->plans.size()) return fail(TDBG_ERR_INVALID, "plan %u not planned", i);
  const tdbg_batch* w = wave_of(const_cast<tdbg_batch*>(b), i);
  if (!w->built || i >= w->info.size()) return fail(TDBG_ERR_INVALID, "plan %u not planned", i);
  *out = w->info[i];
  return TDBG_OK;
}

void tdbg_batch_destroy(tdbg_batch* b) { free_batch(b); }

}  // extern "C"
