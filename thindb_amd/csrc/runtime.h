// runtime.h — host-side internals of libthindb_gpu.so shared by runtime.cpp
// (context, HBM column store, single-query plans) and batch.cpp (batched and
// served execution of many plans).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/thindb_gpu.h"
#include "engine.h"

namespace tdbg {
namespace rt {

// Records `fmt` as the calling thread's last error and returns `code`.
int32_t fail(int32_t code, const char* fmt, ...);
int32_t hip_fail(hipError_t e, const char* what);

#define HIP_TRY(expr, what)                              \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) return ::tdbg::rt::hip_fail(e_, what); \
  } while (0)

int type_bytes(uint32_t t);
inline bool is_int_type(uint32_t t) { return t >= TDBG_BOOLEAN && t <= TDBG_DECIMAL64; }
inline bool is_float_type(uint32_t t) { return t == TDBG_FLOAT || t == TDBG_DOUBLE; }
uint32_t kind_of(uint32_t t);
inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }
inline uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

constexpr uint32_t kTileRows = 16384;        // rows per scan tile
constexpr uint64_t kDenseMax = 1ull << 24;   // key spans up to this take a dense HBM table ...
constexpr uint64_t kDenseFloor = 1ull << 16; // ... always up to this,
constexpr uint64_t kDenseRowsX = 16;         // ... else when at most this many keys per row scanned

// ---- HBM column store ------------------------------------------------------

struct ColumnArena {
  uint32_t type = 0;
  uint32_t elem = 0;
  uint8_t* d_data = nullptr;
  uint32_t* d_valid = nullptr;  // allocated once any block is nullable
  bool has_stats = false;
  int64_t min_i = 0, max_i = 0;
  double min_f = 0, max_f = 0;
};

struct BlockInfo {
  bool staged = false;
  bool has_stats = false;
  tdbg_block_stats stats{};
};

struct RowGroupEntry {
  uint64_t segment_id = 0;
  uint32_t rg_idx = 0;
  uint32_t n_rows = 0;
  uint64_t row_off = 0;
  std::map<uint32_t, BlockInfo> blocks;
};

struct Table {
  uint64_t uid = 0;
  uint64_t cap_rows = 0;
  uint64_t next_row = 0;
  uint64_t reserve = 0;
  uint64_t version = 0;
  std::map<uint32_t, ColumnArena> cols;
  std::vector<RowGroupEntry> rgs;
  std::unordered_map<uint64_t, std::vector<size_t>> rg_index;  // hash(seg, rg) -> entries
};

inline uint64_t rg_hash(uint64_t seg, uint32_t rg) { return seg * 0x9e3779b97f4a7c15ull ^ rg; }

// ---- plans -----------------------------------------------------------------

struct AggOut {
  uint32_t op;          // tdbg_agg_op
  int kagg;             // index into kernel aggs, -1 when served by the row count
  int dslot = -1;       // COUNT(DISTINCT): index into Prepared::distinct
  uint32_t in_type, out_type, scale;
  bool is_float;
  bool nullable;
};

// How prepare() shapes the group table.  S_AUTO lets it choose; the batch
// autotuner passes the strategy it measured fastest for the plan's
// (rows, estimated groups) bucket.
enum Strategy : uint32_t { S_AUTO = 0, S_DIRECT, S_GLOBAL, S_LDS_HASH, S_RADIX };
struct PlanHint {
  uint32_t strategy = S_AUTO;
  uint32_t leaf_bits = 0;       // S_RADIX: log2 keys per leaf (0: largest that fits)
  uint32_t lds_budget = 0;      // LDS bytes per workgroup (0: the device maximum)
  uint32_t max_bits1 = 8;       // S_RADIX: widest level-1 digit
  bool single_level = false;    // S_RADIX: one scatter level only
  bool dense_out = false;       // S_RADIX: leaves write a dense record array
  bool wide_leaves = false;     // S_RADIX: the widest leaf that fits (parallelism comes from elsewhere)
};

struct Prepared {
  uint64_t version = ~0ull;
  uint64_t rows_hint = 0;
  ScanParams sp{};
  OrderParams op{};
  FinalParams fp{};
  uint32_t grid = 0, block = 0;
  std::vector<uint64_t> tmpl;  // one initialised record
  bool empty_domain = false;   // predicates admit no row
  std::vector<uint32_t> used_cols;  // table column per ScanParams col
  bool radix = false;          // high cardinality: radix partitioning (radix.hip)
  RadixParams rx{};
  uint32_t agg_grid = 0;
  // COUNT(DISTINCT) inputs: the table column and the record offset of each
  // group's count (distinct.hip fills it after compaction)
  struct Distinct {
    uint32_t col;
    uint32_t dslot_off;
  };
  std::vector<Distinct> distinct;
  // HAVING evaluated in the radix leaf pass (compacted leaves, every target
  // held in the leaf's LDS table): only passing groups are written
  bool leaf_having = false;
  HavingLeaf hleaf{};
};

struct MergedPred {
  uint32_t col;
  bool is_float;
  bool empty = false;
  int64_t lo_i = INT64_MIN, hi_i = INT64_MAX;
  double lo_f = -HUGE_VAL, hi_f = HUGE_VAL;
  bool lo_incl = true, hi_incl = true;
};

// The row groups one execution scans, as tiles, after zone-map pruning.
struct ScanSet {
  std::vector<MergedPred> ranges;
  std::vector<tdbg_predicate> neqs;
  std::vector<Tile> tiles;
  uint64_t rows = 0;
  uint64_t bytes = 0;  // native-width bytes of the referenced columns
};

// Pinned host buffers for results, recycled: a buffer returns to the pool
// when the last result viewing it (and its plan or batch) lets it go.
struct PinnedPool {
  std::mutex mu;
  std::multimap<uint64_t, uint8_t*> free;  // capacity -> buffer
  ~PinnedPool();
};
struct PinnedBuf {
  std::shared_ptr<uint8_t> p;
  uint64_t cap = 0;
};
int32_t pinned_acquire(const std::shared_ptr<PinnedPool>& pool, uint64_t bytes, PinnedBuf& out);

}  // namespace rt
}  // namespace tdbg

struct tdbg_context {
  int device = 0;
  uint64_t budget = 0;
  std::atomic<uint64_t> in_use{0};
  std::atomic<uint64_t> clock{0};  // table versions: each change takes the next value
  std::shared_mutex mu;
  std::unordered_map<uint64_t, std::unique_ptr<tdbg::rt::Table>> tables;
  int n_cu = 256;
  int max_lds = 160 * 1024;
  // batch autotuner: (rows bucket, groups bucket) -> measured best hint
  std::mutex tune_mu;
  std::map<uint32_t, tdbg::rt::PlanHint> tuned;
  std::shared_ptr<tdbg::rt::PinnedPool> pinned = std::make_shared<tdbg::rt::PinnedPool>();
  // device staging of encoded block bodies (tdbg_column_stage_encoded)
  void* d_stage = nullptr;
  uint64_t d_stage_bytes = 0;
};

struct tdbg_result {
  uint64_t n_rows = 0;
  std::vector<tdbg_column_view> cols;
  std::vector<std::vector<uint8_t>> data;   // owned column copies
  std::vector<std::vector<uint8_t>> valid;
  std::shared_ptr<void> keep;               // host memory the views borrow (batch results)
};

struct tdbg_plan {
  tdbg_context* ctx = nullptr;
  uint64_t table_uid = 0;
  std::vector<tdbg_key_part> keys;
  std::vector<tdbg_agg_spec> aggs;
  std::vector<tdbg_predicate> preds;
  std::vector<tdbg_order_spec> order;
  std::vector<tdbg_having_node> having;  // postfix HAVING nodes, empty = none
  std::vector<tdbg_filter_node> where;   // postfix WHERE tree AND-ed with preds, empty = none
  uint64_t limit = 0, offset = 0, est_groups = 0;
  std::vector<tdbg::rt::AggOut> outs;
  std::mutex mu;
  // tdbg_cancel bumps it; an execution captures it when it is asked for
  // (tdbg_execute's entry, a server ticket's submission) and stops when it
  // has changed since, so a cancel reaches queued and waiting executions
  std::atomic<uint64_t> cancel_gen{0};
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  tdbg::rt::Prepared prep;
  // device buffers, grown on demand
  struct Buf {
    void* p = nullptr;
    uint64_t bytes = 0;
  } state, tiles, err, gids, counts, sel, scratch, out, tmpl, nout, rx_misc, rx_rows1, rx_rows2, rx_leaf, rx_words,
      having_nodes, having_gids, dist_set, dist_counts, having_leaf, dist_entries;
  // prepare() bumps prep_stamp; an execution reuses the preparation it made
  // (exec_prep_*) while the stamp, its scan set and the planner's switches
  // stay the same, and skips re-uploading tiles and the record template
  // while the buffers it uploaded them to (up_*) are the current ones
  uint64_t prep_stamp = 0, exec_prep_stamp = ~0ull, up_stamp = ~0ull;
  std::string exec_prep_env;
  // hash partitions: extra leaf bits after a leaf table overflowed (kept for
  // later executions of the plan)
  uint32_t rx_hash_boost = 0;
  void *up_tiles = nullptr, *up_tmpl = nullptr, *up_hleaf = nullptr;
  // tdbg_execute_noalloc: while set (under `mu`), ensure() carves every
  // buffer from this device scratch region instead of allocating
  struct Arena {
    uint8_t* base = nullptr;
    uint64_t cap = 0, used = 0;
  } arena;
  bool arena_on = false;
  // pinned host staging: inputs (tiles, template) and the result the
  // columns of a returned tdbg_result view
  uint8_t* h_in = nullptr;
  uint64_t h_in_cap = 0;
  tdbg::rt::PinnedBuf h_out;
  tdbg_exec_stats last{};
  // the last execution's scan set, with the table version and row groups it
  // was resolved for
  struct ScanCache {
    bool valid = false;
    const tdbg::rt::Table* table = nullptr;
    uint64_t version = 0;
    bool all = false;
    std::vector<tdbg_rg_ref> rgs;
    tdbg::rt::ScanSet S;
  } scan_cache;
};

namespace tdbg {
namespace rt {

// Plans a batch (and the serving loop's batched lane) can run: no ORDER BY,
// LIMIT, OFFSET or COUNT(DISTINCT).
inline bool plan_batchable(const tdbg_plan* pl) {
  if (!pl->order.empty() || pl->limit || pl->offset) return false;
  for (auto& a : pl->aggs)
    if (a.op == TDBG_AGG_COUNT_DISTINCT) return false;
  return true;
}

int32_t dev_alloc(tdbg_context* ctx, void** p, uint64_t bytes);
void dev_free(tdbg_context* ctx, void* p, uint64_t bytes);
int32_t ensure(tdbg_plan* pl, tdbg_plan::Buf& b, uint64_t bytes);
// Grows a pinned host buffer to at least `bytes`.
int32_t ensure_pinned(uint8_t*& p, uint64_t& cap, uint64_t bytes);
void release(tdbg_plan* pl, tdbg_plan::Buf& b);

// Resolves the row groups a plan scans (all staged ones when rgs is null),
// checks that every referenced block is staged, prunes by zone maps and cuts
// the survivors into tiles.  Caller holds ctx->mu shared.
int32_t collect_scan(tdbg_plan* pl, Table* t, const tdbg_rg_ref* rgs, uint32_t n_rgs, ScanSet& out);

// Specialises the plan to the table's statistics and the scanned row count.
int32_t prepare(tdbg_plan* pl, Table* t, const ScanSet& s, const PlanHint& hint);

// COUNT(DISTINCT) over the compacted groups gids[0, *n_groups) (device count,
// at most max_groups), on the plan's stream.
constexpr uint64_t kDistinctDomainMax = 1ull << 31;
int32_t run_distinct(tdbg_plan* pl, Table* t, const ScanSet& S, const Prepared::Distinct& dd, const uint32_t* gids,
                     const uint64_t* n_groups, uint64_t max_groups);

// A result of n_out rows whose column c sits at host + data_off[c] (validity
// bytes at host + valid_off[c], or ~0 for none): copied out of `host`, or,
// with `keep` set, viewed in place while the result holds `keep`.
int32_t make_result(tdbg_plan* pl, uint64_t n_out, const uint8_t* host, const std::vector<uint64_t>& data_off,
                    const std::vector<uint64_t>& valid_off, std::shared_ptr<void> keep, tdbg_result** out);

// One execution of a plan (tdbg_execute): the result's columns view the
// plan's pinned result buffer (`views`) or are copied out of it, which leaves
// the buffer free for the plan's next execution (the serving loop).
int32_t execute_plan(tdbg_plan* pl, const tdbg_rg_ref* rgs, uint32_t n_rgs, bool views, tdbg_result** out,
                     uint64_t cancel_gen);
inline bool cancelled(const tdbg_plan* pl, uint64_t gen) { return pl->cancel_gen.load() != gen; }

// tdbg_execute_noalloc: the same execution with every device buffer carved
// from `scratch` (device memory of `scratch_bytes`), the plan's staging
// through pageable host memory, the plan-specialised kernel only when
// already compiled, and the result written into the caller's host buffer.
struct NoAllocOut {
  uint8_t* scratch;
  uint64_t scratch_bytes;
  void* out;
  uint64_t out_capacity;
  tdbg_column_view* cols;
  uint32_t n_cols;
  uint64_t* n_rows;
};
int32_t execute_plan_noalloc(tdbg_plan* pl, const tdbg_rg_ref* rgs, uint32_t n_rgs, const NoAllocOut& na);
// Resolves `device`'s copy of the no-alloc scratch array (noalloc.cpp) with
// that device current; idempotent.  Called by tdbg_context_create.
int32_t noalloc_setup(int device);

// A reusable batch for the serving loop (batch.cpp): eager launches, device
// and pinned buffers kept and grown across runs of different plan sets.
int32_t batch_open(tdbg_context* ctx, tdbg_batch** out);
// rgs (optional): per plan, the row groups it scans (null or empty: all staged)
int32_t batch_run_plans(tdbg_batch* b, tdbg_plan* const* plans, uint32_t n,
                        const std::vector<tdbg_rg_ref>* const* rgs = nullptr);
int32_t batch_take_result(tdbg_batch* b, uint32_t i, tdbg_result** out);  // views pooled pinned memory
void batch_close(tdbg_batch* b);

}  // namespace rt
}  // namespace tdbg
