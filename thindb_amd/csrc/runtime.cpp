// runtime.cpp — host side of libthindb_gpu.so: the C-ABI of
// include/thindb_gpu.h over HIP.
//
//  * Context: one GPU, an HBM byte budget (admission: allocations beyond it
//    fail with TDBG_ERR_OOM, the analog of src/memory.zig:31-35), and the
//    resident tables.
//  * HBM column store: per (table_uid, column_idx) one contiguous arena;
//    every staged row group gets one row range [row_off, row_off + n) shared
//    by all its columns, row_off a multiple of kRowAlign, so a scan is a
//    list of row ranges over flat arrays (the decoded-block cache of
//    src/storage/cache.zig:1-8 keyed like cache.Key, :59-70; segments are
//    immutable, so nothing is ever invalidated).
//  * Plan: the validated request; at execution it is specialised to the
//    table's current statistics (key spans -> dense or hashed HBM table,
//    LDS table shape, exact-sum mode) and run as one stream of kernels:
//    record init -> fused scan/filter/aggregate -> compaction -> ORDER BY
//    selection -> finalize -> one device-to-host copy.
#include <functional>

#include "runtime.h"

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>

using namespace tdbg;
using namespace tdbg::rt;

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace tdbg {
namespace rt {

int32_t fail(int32_t code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int32_t hip_fail(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory) return fail(TDBG_ERR_OOM, "%s: %s", what, hipGetErrorString(e));
  return fail(TDBG_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

int type_bytes(uint32_t t) {
  switch (t) {
    case TDBG_BOOLEAN: case TDBG_TINYINT: return 1;
    case TDBG_SMALLINT: return 2;
    case TDBG_INT: case TDBG_DATE: case TDBG_FLOAT: return 4;
    case TDBG_BIGINT: case TDBG_DATETIME: case TDBG_DECIMAL64: case TDBG_DOUBLE: return 8;
    case TDBG_DECIMAL128: return 16;
    default: return 0;
  }
}
uint32_t kind_of(uint32_t t) {
  switch (t) {
    case TDBG_BOOLEAN: return K_U8;
    case TDBG_TINYINT: return K_I8;
    case TDBG_SMALLINT: return K_I16;
    case TDBG_INT: case TDBG_DATE: return K_I32;
    case TDBG_FLOAT: return K_F32;
    case TDBG_DOUBLE: return K_F64;
    default: return K_I64;
  }
}

int32_t dev_alloc(tdbg_context* ctx, void** p, uint64_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  uint64_t now = ctx->in_use.fetch_add(bytes) + bytes;
  if (now > ctx->budget) {
    ctx->in_use.fetch_sub(bytes);
    return fail(TDBG_ERR_OOM, "HBM budget exceeded: %llu of %llu bytes", (unsigned long long)now,
                (unsigned long long)ctx->budget);
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    ctx->in_use.fetch_sub(bytes);
    *p = nullptr;
    (void)hipGetLastError();
    return hip_fail(e, "hipMalloc");
  }
  return TDBG_OK;
}

void dev_free(tdbg_context* ctx, void* p, uint64_t bytes) {
  if (!p) return;
  if (bytes == 0) bytes = 256;
  (void)hipFree(p);
  ctx->in_use.fetch_sub(bytes);
}

int32_t ensure(tdbg_plan* pl, tdbg_plan::Buf& b, uint64_t bytes) {
  if (b.bytes >= bytes && b.p) return TDBG_OK;
  if (pl->arena_on) {  // tdbg_execute_noalloc: carve from the scratch region, never allocate
    tdbg_plan::Arena& A = pl->arena;
    const uint64_t at = (A.used + 255) & ~255ull;
    const uint64_t want = std::max<uint64_t>(bytes, 256);
    if (at > A.cap || want > A.cap - at)
      return fail(TDBG_ERR_UNSUPPORTED, "execution needs more than the %llu-byte no-alloc scratch array",
                  (unsigned long long)A.cap);
    b.p = A.base + at;
    b.bytes = want;
    A.used = at + want;
    return TDBG_OK;
  }
  dev_free(pl->ctx, b.p, b.bytes);
  b.p = nullptr;
  b.bytes = 0;
  uint64_t want = std::max<uint64_t>(bytes, 256);
  int32_t rc = dev_alloc(pl->ctx, &b.p, want);
  if (rc) return rc;
  b.bytes = want;
  return TDBG_OK;
}

void release(tdbg_plan* pl, tdbg_plan::Buf& b) {
  dev_free(pl->ctx, b.p, b.bytes);
  b.p = nullptr;
  b.bytes = 0;
}

PinnedPool::~PinnedPool() {
  for (auto& kv : free) (void)hipHostFree(kv.second);
}

int32_t pinned_acquire(const std::shared_ptr<PinnedPool>& pool, uint64_t bytes, PinnedBuf& out) {
  uint8_t* p = nullptr;
  uint64_t cap = 0;
  {
    std::lock_guard<std::mutex> lk(pool->mu);
    auto it = pool->free.lower_bound(bytes);
    if (it != pool->free.end() && it->first <= 4 * bytes + (1u << 20)) {  // not a far larger one
      cap = it->first;
      p = it->second;
      pool->free.erase(it);
    }
  }
  if (!p) {
    cap = std::max<uint64_t>(bytes, 4096);
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&p), cap, hipHostMallocDefault), "pinned result buffer");
  }
  std::shared_ptr<PinnedPool> keep = pool;
  out.p.reset(p, [keep, cap](uint8_t* q) {
    std::lock_guard<std::mutex> lk(keep->mu);
    keep->free.emplace(cap, q);
  });
  out.cap = cap;
  return TDBG_OK;
}

int32_t ensure_pinned(uint8_t*& p, uint64_t& cap, uint64_t bytes) {
  if (p && cap >= bytes) return TDBG_OK;
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  const uint64_t want = std::max<uint64_t>(bytes, 4096);
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&p), want, hipHostMallocDefault), "pinned staging");
  cap = want;
  return TDBG_OK;
}

}  // namespace rt
}  // namespace tdbg

namespace {

uint64_t arena_bytes(uint64_t cap_rows, uint32_t elem) { return cap_rows * elem + 64; }
uint64_t valid_bytes(uint64_t cap_rows) { return cap_rows / 8 + 64; }

void free_table(tdbg_context* ctx, Table* t) {
  for (auto& kv : t->cols) {
    dev_free(ctx, kv.second.d_data, arena_bytes(t->cap_rows, kv.second.elem));
    if (kv.second.d_valid) dev_free(ctx, kv.second.d_valid, valid_bytes(t->cap_rows));
  }
  t->cols.clear();
}

// Grows every arena of `t` to hold `rows` rows (rounded to kRowAlign).
int32_t table_grow(tdbg_context* ctx, Table* t, uint64_t rows) {
  uint64_t cap = std::max<uint64_t>(rows, t->reserve);
  cap = std::max<uint64_t>(cap, t->cap_rows * 2);
  cap = align_up(std::max<uint64_t>(cap, 1), kRowAlign);
  for (auto& kv : t->cols) {
    ColumnArena& a = kv.second;
    void* nd = nullptr;
    int32_t rc = dev_alloc(ctx, &nd, arena_bytes(cap, a.elem));
    if (rc) return rc;
    if (a.d_data && t->next_row) HIP_TRY(hipMemcpy(nd, a.d_data, t->next_row * a.elem, hipMemcpyDeviceToDevice), "grow copy");
    dev_free(ctx, a.d_data, arena_bytes(t->cap_rows, a.elem));
    a.d_data = static_cast<uint8_t*>(nd);
    if (a.d_valid) {
      void* nv = nullptr;
      rc = dev_alloc(ctx, &nv, valid_bytes(cap));
      if (rc) return rc;
      HIP_TRY(hipMemset(nv, 0xff, valid_bytes(cap)), "grow validity");
      if (t->next_row) HIP_TRY(hipMemcpy(nv, a.d_valid, t->next_row / 8 + 1, hipMemcpyDeviceToDevice), "grow validity copy");
      dev_free(ctx, a.d_valid, valid_bytes(t->cap_rows));
      a.d_valid = static_cast<uint32_t*>(nv);
    }
  }
  t->cap_rows = cap;
  return TDBG_OK;
}

void host_stats(uint32_t type, const void* v, const uint8_t* valid, uint64_t n, tdbg_block_stats* s) {
  s->has_minmax = 0;
  bool first = true;
  for (uint64_t r = 0; r < n; r++) {
    if (valid && !((valid[r >> 3] >> (r & 7)) & 1)) continue;
    if (is_float_type(type)) {
      double x = type == TDBG_FLOAT ? (double)static_cast<const float*>(v)[r] : static_cast<const double*>(v)[r];
      if (x != x) continue;
      if (first) { s->min_f = s->max_f = x; first = false; }
      s->min_f = std::min(s->min_f, x);
      s->max_f = std::max(s->max_f, x);
    } else {
      int64_t x;
      switch (type) {
        case TDBG_BOOLEAN: x = static_cast<const uint8_t*>(v)[r]; break;
        case TDBG_TINYINT: x = static_cast<const int8_t*>(v)[r]; break;
        case TDBG_SMALLINT: x = static_cast<const int16_t*>(v)[r]; break;
        case TDBG_INT: case TDBG_DATE: x = static_cast<const int32_t*>(v)[r]; break;
        default: x = static_cast<const int64_t*>(v)[r]; break;
      }
      if (first) { s->min_i = s->max_i = x; first = false; }
      s->min_i = std::min(s->min_i, x);
      s->max_i = std::max(s->max_i, x);
    }
  }
  s->has_minmax = first ? 0 : 1;
}

}  // namespace

namespace {

// Validates the request; TDBG_ERR_UNSUPPORTED = decline to the CPU path
// (the gates of src/exec/v2_lowcard_group.zig:234-373 and
// src/exec/v2_shape_group_topn.zig:2051-2250 that apply here).
int32_t validate(tdbg_plan* pl) {
  if (pl->keys.empty() || pl->keys.size() > (size_t)kMaxParts) return fail(TDBG_ERR_UNSUPPORTED, "1..8 group keys");
  if (pl->aggs.empty() || pl->aggs.size() > (size_t)kMaxAggs) return fail(TDBG_ERR_UNSUPPORTED, "1..16 aggregates");
  for (auto& k : pl->keys)
    if (!is_int_type(k.type)) return fail(TDBG_ERR_UNSUPPORTED, "non-integer group key");
  std::vector<uint32_t> cols;
  auto use = [&](uint32_t c) {
    if (std::find(cols.begin(), cols.end(), c) == cols.end()) cols.push_back(c);
  };
  for (auto& k : pl->keys) use(k.column_idx);
  for (auto& a : pl->aggs) {
    if (a.op > TDBG_AGG_COUNT_DISTINCT) return fail(TDBG_ERR_INVALID, "aggregate op %u", a.op);
    if (a.op == TDBG_AGG_COUNT_STAR) continue;
    if (a.column_idx < 0) return fail(TDBG_ERR_INVALID, "aggregate without a column");
    if (!is_int_type(a.input_type) && !is_float_type(a.input_type))
      return fail(TDBG_ERR_UNSUPPORTED, "aggregate input type %u", a.input_type);
    if ((a.input_type == TDBG_DATE || a.input_type == TDBG_DATETIME) && (a.op == TDBG_AGG_SUM || a.op == TDBG_AGG_AVG))
      return fail(TDBG_ERR_UNSUPPORTED, "temporal SUM/AVG");
    use((uint32_t)a.column_idx);
  }
  for (auto& p : pl->preds) {
    if (p.op > TDBG_GTE) return fail(TDBG_ERR_INVALID, "predicate op %u", p.op);
    if (!is_int_type(p.type) && !is_float_type(p.type)) return fail(TDBG_ERR_UNSUPPORTED, "predicate type");
    use(p.column_idx);
  }
  if (pl->order.size() > 6) return fail(TDBG_ERR_UNSUPPORTED, "more than 6 ORDER BY items");
  for (auto& o : pl->order)
    if (o.target >= pl->keys.size() + pl->aggs.size()) return fail(TDBG_ERR_INVALID, "ORDER BY target");
  // WHERE tree: a well-formed postfix tree of comparisons over integer or
  // float columns and NULL tests
  if (pl->where.size() > TDBG_MAX_WHERE) return fail(TDBG_ERR_UNSUPPORTED, "more than %u WHERE nodes", TDBG_MAX_WHERE);
  {
    uint64_t depth = 0;
    for (size_t i = 0; i < pl->where.size(); i++) {
      const tdbg_filter_node& f = pl->where[i];
      switch (f.kind) {
        case TDBG_HAVING_CMP:
          if (f.op > TDBG_GTE) return fail(TDBG_ERR_INVALID, "WHERE node %zu: op %u", i, f.op);
          if (!is_int_type(f.type) && !is_float_type(f.type)) return fail(TDBG_ERR_UNSUPPORTED, "WHERE node %zu: type", i);
          use(f.column_idx);
          depth++;
          break;
        case TDBG_HAVING_IS_NULL:
        case TDBG_HAVING_IS_NOT_NULL:
          use(f.column_idx);
          depth++;
          break;
        case TDBG_HAVING_NOT:
          if (depth < 1) return fail(TDBG_ERR_INVALID, "WHERE node %zu: NOT of nothing", i);
          break;
        case TDBG_HAVING_AND:
        case TDBG_HAVING_OR:
          if (f.arity > depth) return fail(TDBG_ERR_INVALID, "WHERE node %zu: arity %u over %llu entries", i, f.arity,
                                           (unsigned long long)depth);
          depth = depth - f.arity + 1;
          break;
        default: return fail(TDBG_ERR_INVALID, "WHERE node %zu: kind %u", i, f.kind);
      }
    }
    if (!pl->where.empty() && depth != 1) return fail(TDBG_ERR_INVALID, "WHERE leaves %llu entries", (unsigned long long)depth);
  }
  if (cols.size() > (size_t)kMaxCols) return fail(TDBG_ERR_UNSUPPORTED, "more than %d distinct columns", kMaxCols);
  // HAVING: a well-formed postfix tree over the output columns (every output
  // here is a key or a COUNT/SUM/AVG/MIN/MAX, all of which the reference's
  // havingLeafColSupported accepts, v2_shape_group_topn.zig:2430-2441)
  static_assert(sizeof(HavingNode) == sizeof(tdbg_having_node) && kMaxHaving == TDBG_MAX_HAVING,
                "HavingNode mirrors tdbg_having_node");
  if (pl->having.size() > kMaxHaving) return fail(TDBG_ERR_UNSUPPORTED, "more than %u HAVING nodes", kMaxHaving);
  {
    const size_t n_out = pl->keys.size() + pl->aggs.size();
    uint64_t depth = 0;
    for (size_t i = 0; i < pl->having.size(); i++) {
      const tdbg_having_node& h = pl->having[i];
      switch (h.kind) {
        case TDBG_HAVING_CMP:
          if (h.op > TDBG_GTE) return fail(TDBG_ERR_INVALID, "HAVING node %zu: op %u", i, h.op);
          if (h.lit_float > 1) return fail(TDBG_ERR_INVALID, "HAVING node %zu: lit_float %u", i, h.lit_float);
          [[fallthrough]];
        case TDBG_HAVING_IS_NULL:
        case TDBG_HAVING_IS_NOT_NULL:
          if (h.target >= n_out) return fail(TDBG_ERR_INVALID, "HAVING node %zu: target %u", i, h.target);
          depth++;
          break;
        case TDBG_HAVING_NOT:
          if (depth < 1) return fail(TDBG_ERR_INVALID, "HAVING node %zu: NOT of nothing", i);
          break;
        case TDBG_HAVING_AND:
        case TDBG_HAVING_OR:
          if (h.arity > depth) return fail(TDBG_ERR_INVALID, "HAVING node %zu: arity %u over %llu entries", i, h.arity,
                                           (unsigned long long)depth);
          depth = depth - h.arity + 1;
          break;
        default: return fail(TDBG_ERR_INVALID, "HAVING node %zu: kind %u", i, h.kind);
      }
    }
    if (!pl->having.empty() && depth != 1) return fail(TDBG_ERR_INVALID, "HAVING leaves %llu entries", (unsigned long long)depth);
  }
  // output types: src/exec/aggregate.zig:3700-3735
  for (auto& a : pl->aggs) {
    AggOut o{};
    o.op = a.op;
    o.kagg = -1;
    o.in_type = a.input_type;
    o.scale = a.input_scale;
    o.is_float = is_float_type(a.input_type);
    o.nullable = a.nullable != 0;
    switch (a.op) {
      case TDBG_AGG_COUNT_STAR: case TDBG_AGG_COUNT: case TDBG_AGG_COUNT_DISTINCT: o.out_type = TDBG_BIGINT; break;
      case TDBG_AGG_SUM:
        o.out_type = o.is_float ? TDBG_DOUBLE : (a.input_type == TDBG_DECIMAL64 ? TDBG_DECIMAL128 : TDBG_BIGINT);
        break;
      case TDBG_AGG_AVG: o.out_type = TDBG_DOUBLE; break;
      default: o.out_type = a.input_type; break;
    }
    pl->outs.push_back(o);
  }
  return TDBG_OK;
}

// The WHERE's `col op literal` leaves merged per column into one range term
// (plus NEQ terms).  Returns false when some column's range is empty.
bool merge_preds(const std::vector<tdbg_predicate>& preds, std::vector<MergedPred>& ranges,
                 std::vector<tdbg_predicate>& neqs) {
  for (auto& p : preds) {
    if (p.op == TDBG_NEQ) {
      neqs.push_back(p);
      continue;
    }
    MergedPred* m = nullptr;
    for (auto& r : ranges)
      if (r.col == p.column_idx) m = &r;
    if (!m) {
      ranges.push_back(MergedPred{p.column_idx, is_float_type(p.type)});
      m = &ranges.back();
    }
    if (!m->is_float) {
      const int64_t x = p.i64;
      switch (p.op) {
        case TDBG_EQ: m->lo_i = std::max(m->lo_i, x); m->hi_i = std::min(m->hi_i, x); break;
        case TDBG_LT: if (x == INT64_MIN) m->empty = true; else m->hi_i = std::min(m->hi_i, x - 1); break;
        case TDBG_LTE: m->hi_i = std::min(m->hi_i, x); break;
        case TDBG_GT: if (x == INT64_MAX) m->empty = true; else m->lo_i = std::max(m->lo_i, x + 1); break;
        default: m->lo_i = std::max(m->lo_i, x); break;
      }
      if (m->lo_i > m->hi_i) m->empty = true;
    } else {
      const double x = p.f64;
      if (x != x) { m->empty = true; continue; }
      auto raise_lo = [&](double v, bool incl) {
        if (v > m->lo_f || (v == m->lo_f && m->lo_incl && !incl)) { m->lo_f = v; m->lo_incl = incl; }
      };
      auto lower_hi = [&](double v, bool incl) {
        if (v < m->hi_f || (v == m->hi_f && m->hi_incl && !incl)) { m->hi_f = v; m->hi_incl = incl; }
      };
      switch (p.op) {
        case TDBG_EQ: raise_lo(x, true); lower_hi(x, true); break;
        case TDBG_LT: lower_hi(x, false); break;
        case TDBG_LTE: lower_hi(x, true); break;
        case TDBG_GT: raise_lo(x, false); break;
        default: raise_lo(x, true); break;
      }
      if (m->lo_f > m->hi_f || (m->lo_f == m->hi_f && !(m->lo_incl && m->hi_incl))) m->empty = true;
    }
  }
  for (auto& r : ranges)
    if (r.empty) return false;
  return true;
}

// rowGroupCanMatch (src/exec/scan.zig:3539-3552) over one block's zone map.
bool block_can_match(const MergedPred& m, const BlockInfo& b) {
  if (!b.has_stats || !b.stats.has_minmax) return true;
  if (!m.is_float) return !(b.stats.max_i < m.lo_i || b.stats.min_i > m.hi_i);
  const double lo = b.stats.min_f, hi = b.stats.max_f;
  if (hi < m.lo_f || (hi == m.lo_f && !m.lo_incl)) return false;
  if (lo > m.hi_f || (lo == m.hi_f && !m.hi_incl)) return false;
  return true;
}

}  // namespace

namespace tdbg {
namespace rt {

int32_t collect_scan(tdbg_plan* pl, Table* t, const tdbg_rg_ref* rgs, uint32_t n_rgs, ScanSet& S) {
  S = ScanSet{};
  // the predicate terms and the columns every row group must have staged
  const bool satisfiable = merge_preds(pl->preds, S.ranges, S.neqs);
  std::vector<uint32_t> need_cols;
  for (auto& k : pl->keys) need_cols.push_back(k.column_idx);
  for (auto& a : pl->aggs) if (a.op != TDBG_AGG_COUNT_STAR) need_cols.push_back((uint32_t)a.column_idx);
  for (auto& p : pl->preds) need_cols.push_back(p.column_idx);
  for (auto& f : pl->where)
    if (f.kind <= TDBG_HAVING_IS_NOT_NULL) need_cols.push_back(f.column_idx);
  std::vector<const RowGroupEntry*> scan;
  auto consider = [&](const RowGroupEntry& rg) -> int32_t {
    for (uint32_t c : need_cols) {
      auto bi = rg.blocks.find(c);
      if (bi == rg.blocks.end() || !bi->second.staged)
        return fail(TDBG_ERR_NOT_STAGED, "segment %llu row group %u column %u not staged",
                    (unsigned long long)rg.segment_id, rg.rg_idx, c);
    }
    for (auto& m : S.ranges) {
      auto bi = rg.blocks.find(m.col);
      if (!block_can_match(m, bi->second)) return TDBG_OK;  // zone map prunes it
    }
    scan.push_back(&rg);
    return TDBG_OK;
  };
  if (rgs) {
    for (uint32_t i = 0; i < n_rgs; i++) {
      auto b = t->rg_index.find(rg_hash(rgs[i].segment_id, rgs[i].row_group_idx));
      const RowGroupEntry* found = nullptr;
      if (b != t->rg_index.end())
        for (size_t j : b->second)
          if (t->rgs[j].segment_id == rgs[i].segment_id && t->rgs[j].rg_idx == rgs[i].row_group_idx) found = &t->rgs[j];
      if (!found) return fail(TDBG_ERR_NOT_STAGED, "segment %llu row group %u not staged",
                              (unsigned long long)rgs[i].segment_id, rgs[i].row_group_idx);
      int32_t rc = consider(*found);
      if (rc) return rc;
    }
  } else {
    for (auto& rg : t->rgs) {
      int32_t rc = consider(rg);
      if (rc) return rc;
    }
  }
  if (!satisfiable) return TDBG_OK;
  for (const RowGroupEntry* rg : scan) {
    for (uint32_t o = 0; o < rg->n_rows; o += kTileRows) {
      Tile tl{};
      tl.row0 = rg->row_off + o;
      tl.n = std::min<uint32_t>(kTileRows, rg->n_rows - o);
      S.tiles.push_back(tl);
    }
    S.rows += rg->n_rows;
  }
  std::vector<uint32_t> cols;
  for (uint32_t c : need_cols)
    if (std::find(cols.begin(), cols.end(), c) == cols.end()) cols.push_back(c);
  for (uint32_t c : cols) {
    auto it = t->cols.find(c);
    if (it != t->cols.end()) S.bytes += S.rows * (uint64_t)it->second.elem;
  }
  return TDBG_OK;
}

int32_t prepare(tdbg_plan* pl, Table* t, const ScanSet& S, const PlanHint& hint) {
  pl->prep_stamp++;  // what an execution prepared and uploaded is stale from here on
  const uint64_t n_rows_scan = S.rows;
  const uint32_t n_tiles = (uint32_t)S.tiles.size();
  const std::vector<MergedPred>& ranges = S.ranges;
  const std::vector<tdbg_predicate>& neqs = S.neqs;
  Prepared& P = pl->prep;
  P = Prepared{};
  ScanParams& sp = P.sp;
  auto col_slot = [&](uint32_t table_col) -> uint32_t {
    for (uint32_t i = 0; i < P.used_cols.size(); i++)
      if (P.used_cols[i] == table_col) return i;
    P.used_cols.push_back(table_col);
    return (uint32_t)P.used_cols.size() - 1;
  };
  // predicates
  for (auto& r : ranges) {
    PredDesc d{};
    d.kind = r.is_float ? P_RANGE_F : P_RANGE_I;
    d.col = col_slot(r.col);
    d.lo_i = r.lo_i; d.hi_i = r.hi_i;
    d.lo_f = r.lo_f; d.hi_f = r.hi_f;
    d.lo_incl = r.lo_incl; d.hi_incl = r.hi_incl;
    sp.preds[sp.n_preds++] = d;
  }
  for (auto& q : neqs) {
    PredDesc d{};
    d.kind = is_float_type(q.type) ? P_NEQ_F : P_NEQ_I;
    d.col = col_slot(q.column_idx);
    d.lo_i = q.i64;
    d.lo_f = q.f64;
    sp.preds[sp.n_preds++] = d;
  }
  if (sp.n_preds > (uint32_t)kMaxPreds) return fail(TDBG_ERR_UNSUPPORTED, "too many predicate terms");
  // the WHERE tree: each comparison as one range or NEQ term of its own
  sp.n_filt = 0;
  for (const tdbg_filter_node& f : pl->where) {
    FilterNode& n = sp.filt[sp.n_filt++];
    n = FilterNode{};
    n.kind = f.kind;
    n.arity = f.arity;
    if (f.kind > TDBG_HAVING_IS_NOT_NULL) continue;
    auto it = t->cols.find(f.column_idx);
    if (it == t->cols.end()) return fail(TDBG_ERR_NOT_STAGED, "WHERE column %u not staged", f.column_idx);
    n.pd.col = col_slot(f.column_idx);
    if (f.kind != TDBG_HAVING_CMP) continue;
    if (it->second.type != f.type) return fail(TDBG_ERR_INVALID, "WHERE column %u type mismatch", f.column_idx);
    tdbg_predicate leaf{f.column_idx, f.type, f.op, 0u, f.i64, f.f64};
    std::vector<MergedPred> one;
    std::vector<tdbg_predicate> neq;
    const bool sat = merge_preds({leaf}, one, neq);
    if (!neq.empty()) {
      n.pd.kind = is_float_type(f.type) ? P_NEQ_F : P_NEQ_I;
      n.pd.lo_i = f.i64;
      n.pd.lo_f = f.f64;
    } else {
      const MergedPred& m = one[0];
      n.pd.kind = m.is_float ? P_RANGE_F : P_RANGE_I;
      n.pd.lo_i = m.lo_i; n.pd.hi_i = m.hi_i;
      n.pd.lo_f = m.lo_f; n.pd.hi_f = m.hi_f;
      n.pd.lo_incl = m.lo_incl; n.pd.hi_incl = m.hi_incl;
      if (!sat) {  // a comparison no value meets
        n.pd.lo_i = 1; n.pd.hi_i = 0;
        n.pd.lo_f = 1.0; n.pd.hi_f = 0.0;
      }
    }
  }
  // key parts: relative values in a mixed-radix index, or bit fields of a
  // packed key of up to 128 bits when the domain reaches 2^64 (the silo packs
  // keys into at most 128 bits, v2_shape_group_topn.zig:2051-2094)
  unsigned __int128 domain = 1;
  const unsigned __int128 kDomainCap = (unsigned __int128)1 << 100;  // saturates: only compared with 2^64
  uint32_t key_bits = 0, part_bits[kMaxParts] = {};
  for (size_t i = 0; i < pl->keys.size(); i++) {
    auto it = t->cols.find(pl->keys[i].column_idx);
    if (it == t->cols.end()) return fail(TDBG_ERR_NOT_STAGED, "key column %u not staged", pl->keys[i].column_idx);
    const ColumnArena& a = it->second;
    if (a.type != pl->keys[i].type) return fail(TDBG_ERR_INVALID, "key column %u type mismatch", pl->keys[i].column_idx);
    PartDesc d{};
    d.col = col_slot(pl->keys[i].column_idx);
    d.base = a.has_stats ? a.min_i : 0;
    unsigned __int128 span = a.has_stats ? (unsigned __int128)((uint64_t)a.max_i - (uint64_t)a.min_i) + 1 : 1;
    if (a.d_valid) {
      // NULL takes relative value 0, the values 1.. (packValidity: the NULL
      // group stays apart from every value, v2_lowcard_group.zig:44-49)
      d.nullable = 1;
      d.base = (int64_t)((uint64_t)d.base - 1);
      span = a.has_stats ? span + 1 : 1;
    }
    if (span > ((unsigned __int128)1 << 64))
      return fail(TDBG_ERR_UNSUPPORTED, "nullable key part over all 2^64 values (65-bit field)");
    d.span = (uint64_t)span;  // 0: all 2^64 values
    uint32_t w = 0;
    while (w < 64 && ((unsigned __int128)1 << w) < span) w++;
    part_bits[i] = w;
    key_bits += w;
    d.stride = domain < ((unsigned __int128)1 << 64) ? (uint64_t)domain : 0;
    domain = domain > kDomainCap / span ? kDomainCap : domain * span;
    sp.parts[sp.n_parts++] = d;
  }
  if (key_bits > 128) return fail(TDBG_ERR_UNSUPPORTED, "group key packs into %u bits (more than 128)", key_bits);
  // a domain of 2^62 or more has only the hash-partitioned path; of 2^64 or
  // more it needs the packed 128-bit key (G_WIDE)
  const bool big_domain = domain >= ((unsigned __int128)1 << 62);
  const bool packed_key = domain >= ((unsigned __int128)1 << 64);
  if (packed_key) {
    for (uint32_t i = 0, off = 0; i < sp.n_parts; i++) {
      sp.parts[i].bit_off = off;
      sp.parts[i].bit_width = part_bits[i];
      sp.parts[i].stride = 1;
      off += part_bits[i];
    }
  }
  const uint64_t dom = domain > (unsigned __int128)~0ull ? ~0ull : (uint64_t)domain;
  // aggregates (kernel states shared by aggs of the same class and column)
  struct KAgg { uint32_t kind; uint32_t col; bool nullable; };
  std::vector<KAgg> kaggs;
  for (size_t i = 0; i < pl->aggs.size(); i++) {
    const tdbg_agg_spec& a = pl->aggs[i];
    AggOut& o = pl->outs[i];
    o.kagg = -1;
    if (a.op == TDBG_AGG_COUNT_STAR) continue;
    auto it = t->cols.find((uint32_t)a.column_idx);
    if (it == t->cols.end()) return fail(TDBG_ERR_NOT_STAGED, "column %d not staged", a.column_idx);
    if (it->second.type != a.input_type) return fail(TDBG_ERR_INVALID, "column %d type mismatch", a.column_idx);
    const bool nullable = it->second.d_valid != nullptr;
    o.nullable = nullable;
    if (a.op == TDBG_AGG_COUNT_DISTINCT) {  // a count slot of its own, filled after compaction
      o.nullable = false;
      o.dslot = (int)P.distinct.size();
      P.distinct.push_back({(uint32_t)a.column_idx, 0u});
      continue;
    }
    if (a.op == TDBG_AGG_COUNT && !nullable) continue;  // = COUNT(*)
    uint32_t kind;
    switch (a.op) {
      case TDBG_AGG_COUNT: kind = A_COUNT_COL; break;
      case TDBG_AGG_SUM: case TDBG_AGG_AVG: kind = o.is_float ? A_SUM_F : A_SUM_I; break;
      case TDBG_AGG_MIN: kind = o.is_float ? A_MIN_F : A_MIN_I; break;
      default: kind = o.is_float ? A_MAX_F : A_MAX_I; break;
    }
    int found = -1;
    for (size_t k = 0; k < kaggs.size(); k++)
      if (kaggs[k].kind == kind && kaggs[k].col == (uint32_t)a.column_idx) found = (int)k;
    if (found < 0) {
      kaggs.push_back({kind, (uint32_t)a.column_idx, nullable});
      found = (int)kaggs.size() - 1;
    }
    o.kagg = found;
  }
  // HBM record layout: dense (a record per relative key) when the key span is
  // small or well filled by the rows scanned, else a hash table sized for the
  // groups: a dense table over a span far wider than the rows is mostly
  // empty records that the result pass still reads
  const bool dense = (dom <= kDenseMax && (dom <= kDenseFloor || dom <= kDenseRowsX * std::max<uint64_t>(n_rows_scan, 1))) ||
                     (pl->est_groups && dom <= 4 * pl->est_groups);
  sp.g_mode = dense ? G_DENSE : G_HASH;
  uint32_t off = 0;
  if (!dense) { sp.g_key_off = off; off += 8; }
  sp.g_count_off = off; off += 8;
  std::vector<uint64_t> tmpl;
  auto tmpl_put = [&](uint32_t at, uint64_t v) {
    if (tmpl.size() < at / 8 + 1) tmpl.resize(at / 8 + 1, 0);
    tmpl[at / 8] = v;
  };
  if (!dense) tmpl_put(sp.g_key_off, kEmptyKey);
  tmpl_put(sp.g_count_off, 0);
  sp.n_aggs = (uint32_t)kaggs.size();
  for (size_t k = 0; k < kaggs.size(); k++) {
    AggDesc& d = sp.aggs[k];
    d = AggDesc{};
    d.kind = kaggs[k].kind;
    d.col = col_slot(kaggs[k].col);
    d.nullable = kaggs[k].nullable && d.kind != A_COUNT_COL;
    d.g_off = off;
    uint64_t init = 0;
    if (d.kind == A_MIN_I || d.kind == A_MIN_F) init = 0x7fffffffffffffffull;
    if (d.kind == A_MAX_I || d.kind == A_MAX_F) init = 0x8000000000000000ull;
    // an integer SUM state read only as a wrapped BIGINT (no AVG, no DECIMAL
    // SUM over it) needs just its low word: single wrapping HBM adds, and an
    // 8-byte state instead of the i128
    if (d.kind == A_SUM_I) {
      bool lo = true;
      for (auto& o : pl->outs)
        if (o.kagg == (int)k && (o.op != TDBG_AGG_SUM || o.out_type == TDBG_DECIMAL128)) lo = false;
      d.lo_only = lo ? 1u : 0u;
    }
    tmpl_put(off, init);
    off += 8;
    if (d.kind == A_SUM_I && !d.lo_only) { tmpl_put(off, 0); off += 8; }
    d.g_ns_off = ~0u;
    if (d.nullable) { d.g_ns_off = off; tmpl_put(off, 0); off += 8; }
  }
  for (auto& d : P.distinct) {
    d.dslot_off = off;
    tmpl_put(off, 0);
    off += 8;
  }
  sp.g_rec = std::max<uint32_t>(8, off);
  tmpl.resize(sp.g_rec / 8, 0);
  P.tmpl = tmpl;
  if (P.used_cols.size() > (size_t)kMaxCols) return fail(TDBG_ERR_UNSUPPORTED, "more than %d distinct columns", kMaxCols);
  // columns
  sp.n_cols = (uint32_t)P.used_cols.size();
  for (uint32_t i = 0; i < sp.n_cols; i++) {
    auto it = t->cols.find(P.used_cols[i]);
    if (it == t->cols.end()) return fail(TDBG_ERR_NOT_STAGED, "column %u not staged", P.used_cols[i]);
    sp.cols[i].data = it->second.d_data;
    sp.cols[i].valid = it->second.d_valid;
    sp.cols[i].kind = kind_of(it->second.type);
  }
  // estimated groups
  uint64_t est = pl->est_groups ? std::min<uint64_t>(pl->est_groups, dom) : std::min<uint64_t>(dom, std::max<uint64_t>(n_rows_scan, 1));
  if (dense) {
    sp.g_cap = dom;
  } else {
    sp.g_cap = next_pow2(std::max<uint64_t>(1024, est + est / 2 + 1));
  }
  // launch shape and LDS table
  const uint32_t n_cu = (uint32_t)std::max(pl->ctx->n_cu, 8);
  uint32_t budget = std::min<uint32_t>(kLdsBudget, (uint32_t)pl->ctx->max_lds - 1024);
  if (hint.lds_budget) budget = std::min(budget, hint.lds_budget);
  const uint32_t key_bytes = dom < 0xffffffffull ? 4 : 8;
  // LDS bytes of one slot when a workgroup folds `rows_per_wg` rows: an
  // integer SUM keeps one i64 partial when rows_per_wg * max|v| < 2^62,
  // else split 32-bit halves.
  auto slot_bytes_for = [&](uint64_t rows_per_wg) -> uint32_t {
    uint32_t sb = 4;  // u32 row count
    for (uint32_t k = 0; k < sp.n_aggs; k++) {
      AggDesc& d = sp.aggs[k];
      sb += d.kind == A_COUNT_COL ? 4 : 8;
      if (d.nullable) sb += 4;
      if (d.kind != A_SUM_I) continue;
      const ColumnArena& a = t->cols[kaggs[k].col];
      const double m = a.has_stats ? std::max(std::fabs((double)a.min_i), std::fabs((double)a.max_i)) : 9.3e18;
      d.split = ((double)rows_per_wg * m >= 4.0e18) ? 1u : 0u;
      if (d.split) sb += 8;
    }
    return sb;
  };
  uint32_t lds_cap = 0, parts = 1;
  sp.lds_mode = LDS_NONE;
  const uint32_t strat = hint.strategy;
  if (dense && (strat == S_AUTO || strat == S_DIRECT)) {
    // a direct table, split over `parts` workgroups per tile range when the
    // key span does not fit one workgroup's LDS
    const uint32_t max_parts = strat == S_DIRECT ? 64u : 8u;
    for (uint32_t p = 1; p <= max_parts; p *= 2) {
      const uint32_t groups = std::max<uint32_t>(8, n_cu / p);
      const uint64_t rows_per_wg = ((uint64_t)n_tiles + groups - 1) / groups * kTileRows;
      const uint32_t sb = slot_bytes_for(rows_per_wg);
      const uint64_t cap = (dom + p - 1) / p;
      if (cap * sb + 8 * (2 * sp.n_aggs + 3) + 16 <= budget) {  // + alignment of each LDS array
        sp.lds_mode = LDS_DIRECT;
        lds_cap = (uint32_t)cap;
        parts = p;
        break;
      }
    }
  }
  if (sp.lds_mode == LDS_NONE && (strat == S_AUTO || strat == S_LDS_HASH)) {
    const uint32_t sb = slot_bytes_for(((uint64_t)n_tiles + n_cu - 1) / n_cu * kTileRows);
    // every LDS array starts 8-byte aligned and the layout ends 16-byte aligned
    const uint32_t pad = 8 * (2 * sp.n_aggs + 3) + 16;
    const uint32_t cap = (budget - pad) / (sb + key_bytes);
    if ((strat == S_LDS_HASH || est <= 8ull * cap) && cap >= 256) {
      sp.lds_mode = LDS_HASH;
      lds_cap = cap;
    }
  }
  // High cardinality over a key domain below 2^32: radix partitioning by the
  // key's high bits until a leaf's key range fits an LDS direct table.
  // THINDB_GPU_RADIX=0 never takes this path, =1 takes it whenever the
  // domain allows (tests use it to cover the partition passes on small tables).
  // The partitioned row (RowLayout) and the scatter tiles: the u32 relative
  // key or hash, the validity word, the aggregate inputs at native width
  // (aggregates over one column share it), then the key words of hash
  // partitions (last, 8-byte aligned, so narrow inputs fill the gap after
  // the hash: a hash, an INT input and a 128-bit key make 24 bytes).
  auto lay_out_rows = [&](RadixParams& R, uint32_t key_words) {
    R.row.flags_off = ~0u;
    uint32_t ro = 4;
    for (uint32_t k = 0; k < sp.n_aggs; k++)
      if (sp.cols[sp.aggs[k].col].valid) R.row.flags_off = 0;
    if (R.row.flags_off == 0) { R.row.flags_off = ro; ro += 4; }
    for (uint32_t k = 0; k < sp.n_aggs; k++) {
      const AggDesc& d = sp.aggs[k];
      const uint32_t b = d.kind == A_COUNT_COL ? 0u : (uint32_t)t->cols[kaggs[k].col].elem;
      R.row.val_bytes[k] = b;
      int same = -1;
      for (uint32_t j = 0; j < k; j++)
        if (b && sp.aggs[j].col == d.col && R.row.val_bytes[j] == b) same = (int)j;
      if (same >= 0) { R.row.val_off[k] = R.row.val_off[same]; continue; }
      ro = (uint32_t)align_up(ro, std::max<uint32_t>(b, 1));
      R.row.val_off[k] = ro;
      ro += b;
    }
    if (key_words) {
      ro = (uint32_t)align_up(ro, 8);
      R.row.key_off = ro;
      R.row.key_words = key_words;
      ro += 8 * key_words;
    }
    R.row.stride = (uint32_t)align_up(ro, 8);
    // staged scatter tile: 24 KiB of digit tables + tile_rows rows in LDS
    R.tile_rows = std::min<uint32_t>(4096, (budget - 24 * 1024) / R.row.stride / 256 * 256);
    R.tile_rows2 = std::min<uint32_t>(8192, (budget - 24 * 1024) / R.row.stride / 256 * 256);
  };
  const char* rx_env = getenv("THINDB_GPU_RADIX");
  const bool forced = strat == S_RADIX || (strat == S_AUTO && rx_env && rx_env[0] == '1');
  const bool rx_off = (strat != S_AUTO && strat != S_RADIX) || (strat == S_AUTO && rx_env && rx_env[0] == '0');
  if ((sp.lds_mode == LDS_NONE || forced) && dom <= 0xffffffffull && !rx_off) {
    const uint32_t sb = slot_bytes_for(n_rows_scan);  // a leaf may hold every row
    const uint32_t L = dom > 1 ? 64u - (uint32_t)__builtin_clzll(dom - 1) : 0u;  // key bits
    // two leaf workgroups per CU (THINDB_GPU_RADIX_LEAF_BITS caps the table)
    const uint32_t leaf_budget = std::min<uint32_t>(budget, (uint32_t)pl->ctx->max_lds / 2) - 4096;
    const char* lb_env = getenv("THINDB_GPU_RADIX_LEAF_BITS");
    const uint32_t lb_max = hint.leaf_bits ? hint.leaf_bits : (lb_env ? (uint32_t)atoi(lb_env) : 32u);
    uint32_t fit_bits = 0;  // the widest leaf table that fits
    while (fit_bits < L && fit_bits < lb_max && ((uint64_t)sb << (fit_bits + 1)) <= leaf_budget) fit_bits++;
    // at least 2^10 leaves where the domain allows (several per workgroup of
    // the leaf pass), each of at least 64 keys
    // (a batch runs many plans at once and takes the widest leaf instead)
    const uint32_t leaf_bits =
        hint.wide_leaves ? fit_bits : std::min<uint32_t>(fit_bits, std::max<uint32_t>(6, L > 10 ? L - 10 : 0));
    RadixParams& R = P.rx;
    R = RadixParams{};
    lay_out_rows(R, 0);
    const uint32_t rem = L - leaf_bits;
    // one level up to 2^10 buckets, else 2^max_bits1 x the rest (the first
    // level widened up to 2^10 when the second would exceed 2^10)
    const uint32_t max_bits1 = std::min<uint32_t>(hint.max_bits1 ? hint.max_bits1 : 8u, 10u);
    const uint32_t max_rem = hint.single_level ? 10u : 20u;
    // automatic: at least one row per four keys of the domain on average, so
    // leaves are well filled (sparser domains stay with the HBM hash table)
    const bool fits = fit_bits >= 8 && dom > 65536 && n_rows_scan >= (1u << 20) && dom / 4 <= n_rows_scan;
    if ((forced || fits) && rem <= max_rem && R.tile_rows >= 256) {
      P.radix = true;
      R.leaf_bits = leaf_bits;
      R.bits1 = rem <= 10 ? rem : std::max<uint32_t>(std::min<uint32_t>(max_bits1, rem), rem - 10);
      const char* b1_env = getenv("THINDB_GPU_RADIX_BITS1");  // tests: force a second level
      if (b1_env && !hint.single_level) R.bits1 = std::min<uint32_t>(R.bits1, (uint32_t)atoi(b1_env));
      R.bits2 = rem - R.bits1;
      R.shift2 = leaf_bits;
      R.shift1 = leaf_bits + R.bits2;
      R.n_leaves = 1u << rem;
      sp.lds_mode = LDS_DIRECT;  // one leaf's table
      lds_cap = 1u << leaf_bits;
      parts = 1;
      // a mostly occupied dense domain: leaves write every key of their range
      // into the dense table (compaction then yields key order, no sort)
      if (dense && (hint.dense_out || dom <= 2 * est)) {
        R.dense_out = 1;
      } else {
        // records carry their key (the hash layout): they are emitted
        // compacted, leaf by leaf
        if (dense) {
          sp.g_key_off = 0;
          sp.g_count_off += 8;
          for (uint32_t k = 0; k < sp.n_aggs; k++) {
            sp.aggs[k].g_off += 8;
            if (sp.aggs[k].g_ns_off != ~0u) sp.aggs[k].g_ns_off += 8;
          }
          for (auto& dd : P.distinct) dd.dslot_off += 8;
          sp.g_rec += 8;
          P.tmpl.insert(P.tmpl.begin(), kEmptyKey);
        }
        sp.g_mode = G_HASH;
        // a leaf's records start at its first row index: one per row
        sp.g_cap = std::max<uint64_t>(1, n_rows_scan);
      }
    }
  }
  // Wide or sparse keys: hash partitions (the silo routes rows by hash bits,
  // silo_group_core.zig:5226-5235), leaves folding into LDS hash tables.  A
  // domain of 2^62 or more has no other path; a sparse domain takes it over
  // per-row HBM atomics once the table is large.  THINDB_GPU_RADIX_HASH=1
  // takes it whenever the plan allows, =0 never.
  const char* rh_env = getenv("THINDB_GPU_RADIX_HASH");
  const bool h_forced = rh_env && rh_env[0] == '1';
  const bool h_off = rh_env && rh_env[0] == '0';
  if (!P.radix && !h_off && (strat == S_AUTO || strat == S_RADIX) && P.distinct.empty() &&
      (big_domain || h_forced || (sp.lds_mode == LDS_NONE && !dense && !rx_off && n_rows_scan >= (1u << 20)))) {
    RadixParams& R = P.rx;
    R = RadixParams{};
    const uint32_t kw = packed_key ? 2u : 1u;
    lay_out_rows(R, kw);
    // a slot: tag, key words, then the aggregate states; two leaf
    // workgroups per CU
    const uint32_t sb = slot_bytes_for(n_rows_scan) + 8 + 8 * kw;  // + tag, list entry, key words
    const char* wg_env = getenv("THINDB_GPU_RADIX_HASH_WGS");  // leaf workgroups per CU (tuning)
    const uint32_t wgs = std::max<uint32_t>(1, std::min<uint32_t>(4, wg_env ? (uint32_t)atoi(wg_env) : 2u));
    const uint32_t leaf_budget = std::min<uint32_t>(budget, (uint32_t)pl->ctx->max_lds / wgs) - 4096;
    const uint32_t pad = 8 * (2 * sp.n_aggs + 5) + 16;
    const uint32_t cap_max = std::min<uint32_t>((leaf_budget - pad) / sb / 64 * 64, 8192);
    // leaves for the estimated groups at most ~45% of a table on average
    // (the fullest of them then stays well inside it); more after a table
    // overflowed
    const uint64_t G = std::max<uint64_t>(1, pl->est_groups ? std::min<uint64_t>(pl->est_groups, dom)
                                                            : std::min<uint64_t>(n_rows_scan, dom));
    const char* fill_env = getenv("THINDB_GPU_RADIX_HASH_FILL");  // percent of a table (tuning)
    const uint64_t fill = fill_env ? std::max(10, std::min(90, atoi(fill_env))) : 45;
    uint32_t rem = 8;
    while (rem < 20 && (G >> rem) > (uint64_t)cap_max * fill / 100) rem++;
    rem = std::min<uint32_t>(20, rem + pl->rx_hash_boost);
    const uint64_t per_leaf = (G + (1ull << rem) - 1) >> rem;
    if (cap_max >= 256 && R.tile_rows >= 256) {
      P.radix = true;
      const uint32_t cap = (uint32_t)std::min<uint64_t>(cap_max, std::max<uint64_t>(256, align_up(2 * per_leaf + 128, 64)));
      const uint32_t max_bits1 = std::min<uint32_t>(hint.max_bits1 ? hint.max_bits1 : 8u, 10u);
      R.hashed = 1;
      // tests narrow the tags to force collisions (THINDB_GPU_RADIX_HASH_TAG_BITS)
      // or run the exact fold alone (THINDB_GPU_RADIX_HASH_EXACT=1)
      const char* tb_env = getenv("THINDB_GPU_RADIX_HASH_TAG_BITS");
      const uint32_t tb = tb_env ? std::min<uint32_t>(32, std::max<uint32_t>(1, (uint32_t)atoi(tb_env))) : 32u;
      R.tag_mask = tb >= 32 ? ~0u : (1u << tb) - 1u;
      const char* ex_env = getenv("THINDB_GPU_RADIX_HASH_EXACT");
      R.hash_exact = ex_env && ex_env[0] == '1' ? 1u : 0u;
      R.bits1 = rem <= 10 ? rem : std::max<uint32_t>(std::min<uint32_t>(max_bits1, rem), rem - 10);
      R.bits2 = rem - R.bits1;
      R.leaf_bits = 32 - rem;  // hash bits below the digits
      R.shift1 = 32 - R.bits1;
      R.shift2 = 32 - rem;
      R.n_leaves = 1u << rem;
      sp.lds_mode = LDS_HASH;
      lds_cap = cap;
      parts = 1;
      // keyed records, compacted leaf by leaf from each leaf's first row
      const uint32_t had = sp.g_mode == G_DENSE ? 0u : 8u;  // key bytes the layout holds
      const uint32_t shift = 8 * kw - had;
      sp.g_key_off = 0;
      sp.g_count_off += shift;
      for (uint32_t k = 0; k < sp.n_aggs; k++) {
        sp.aggs[k].g_off += shift;
        if (sp.aggs[k].g_ns_off != ~0u) sp.aggs[k].g_ns_off += shift;
      }
      for (auto& dd : P.distinct) dd.dslot_off += shift;
      sp.g_rec += shift;
      P.tmpl.insert(P.tmpl.begin(), shift / 8, kEmptyKey);
      sp.g_mode = packed_key ? G_WIDE : G_HASH;
      sp.g_cap = std::max<uint64_t>(1, n_rows_scan);
    }
  }
  if (big_domain && !P.radix)
    return fail(TDBG_ERR_UNSUPPORTED, "a key domain of 2^62 or more values runs only on hash partitions%s",
                P.distinct.empty() ? "" : " (not with COUNT(DISTINCT))");
  sp.lds_cap = lds_cap;
  sp.lds_parts = parts;
  sp.lds_key64 = key_bytes == 8 || P.rx.hashed;
  uint32_t lo = 0;
  auto lds_take = [&](uint32_t bytes) {
    lo = (uint32_t)align_up(lo, 8);
    uint32_t at = lo;
    lo += bytes;
    return at;
  };
  if (sp.lds_mode != LDS_NONE) {
    if (sp.lds_mode == LDS_HASH) sp.lds_keys_off = lds_take(lds_cap * (P.rx.hashed ? 8 * P.rx.row.key_words : key_bytes));
    if (P.rx.hashed) {
      P.rx.tag_off = lds_take(lds_cap * 4);
      P.rx.list_off = lds_take(lds_cap * 4);
    }
    for (uint32_t k = 0; k < sp.n_aggs; k++) {
      AggDesc& d = sp.aggs[k];
      if (d.kind == A_COUNT_COL) continue;
      d.lds_off = lds_take(lds_cap * 8 * ((d.kind == A_SUM_I && d.split) ? 2 : 1));
    }
    sp.lds_count_off = lds_take(lds_cap * 4);
    for (uint32_t k = 0; k < sp.n_aggs; k++) {
      AggDesc& d = sp.aggs[k];
      if (d.kind == A_COUNT_COL) d.lds_off = lds_take(lds_cap * 4);
      if (d.nullable) d.lds_ns_off = lds_take(lds_cap * 4);
    }
    sp.lds_bytes = (uint32_t)align_up(lo, 16);
    if (sp.lds_bytes > budget) return fail(TDBG_ERR_DEVICE, "LDS layout %u bytes over budget", sp.lds_bytes);
    if (parts > 1) {
      // one 1024-thread workgroup per CU; grid/parts a multiple of 8 keeps
      // a tile range's partners on one XCD
      P.block = 1024;
      P.grid = std::max<uint32_t>(8 * parts, n_cu / (8 * parts) * (8 * parts));
    } else {
      const uint32_t per_cu = std::max<uint32_t>(1, std::min<uint32_t>(2, (uint32_t)pl->ctx->max_lds / std::max<uint32_t>(sp.lds_bytes, 1)));
      P.block = per_cu >= 2 ? 512 : 1024;
      P.grid = (uint32_t)std::min<uint64_t>((uint64_t)n_cu * per_cu, std::max<uint32_t>(n_tiles, 1));
    }
  } else {
    P.block = 256;
    P.grid = (uint32_t)std::min<uint64_t>((uint64_t)n_cu * 8, std::max<uint32_t>(n_tiles, 1) * 4ull);
  }
  if (P.radix) {
    // partition passes: one 1024-thread workgroup per CU over contiguous tile
    // shares; leaf aggregation: persistent 512-thread workgroups, as many
    // per CU as their LDS tables allow
    P.block = 1024;
    P.grid = std::max<uint32_t>(1, std::min<uint32_t>(n_cu, n_tiles));
    const uint32_t per_cu = std::max<uint32_t>(1, std::min<uint32_t>(4, (uint32_t)pl->ctx->max_lds / (sp.lds_bytes + 4096)));
    P.agg_grid = std::min<uint32_t>(P.rx.n_leaves, n_cu * per_cu);
  }
  // order words and output writers
  OrderParams& op = P.op;
  FinalParams& fp = P.fp;
  op.n_parts = fp.n_parts = sp.n_parts;
  for (uint32_t i = 0; i < sp.n_parts; i++) op.parts[i] = fp.parts[i] = sp.parts[i];
  op.g_rec = fp.g_rec = sp.g_rec;
  op.g_count_off = fp.g_count_off = sp.g_count_off;
  op.g_key_off = fp.g_key_off = sp.g_key_off;
  op.g_mode = fp.g_mode = sp.g_mode;
  auto add_comp = [&](uint32_t src, uint32_t idx, uint32_t at, uint32_t ns_off, bool desc, uint32_t aux) -> bool {
    if (op.n_comps >= (uint32_t)kMaxComps) return false;
    CompDesc c{};
    c.src = src; c.idx = idx; c.off = at; c.ns_off = ns_off; c.desc = desc ? 1 : 0; c.aux = aux;
    op.comps[op.n_comps++] = c;
    return true;
  };
  // Every group's SUM over the column fits in 64 bits when the largest
  // |value| times the table's rows does (every staged block's zone map known)
  auto sum_fits_i64 = [&](uint32_t col) -> bool {
    const auto it = t->cols.find(col);
    if (it == t->cols.end() || !it->second.has_stats) return false;
    for (const auto& rg : t->rgs) {
      const auto b = rg.blocks.find(col);
      if (b != rg.blocks.end() && b->second.staged && !b->second.has_stats) return false;
    }
    auto mag = [](int64_t v) { return v < 0 ? (unsigned __int128)(-(__int128)v) : (unsigned __int128)v; };
    const unsigned __int128 m = std::max(mag(it->second.min_i), mag(it->second.max_i));
    return m * (unsigned __int128)t->next_row < ((unsigned __int128)1 << 63);
  };
  bool comps_fit = true;
  for (auto& o : pl->order) {
    const uint32_t nk = (uint32_t)pl->keys.size();
    auto add = [&](uint32_t src, uint32_t at, uint32_t ns_off, uint32_t aux = 0) {
      comps_fit = add_comp(src, src == C_KEY ? o.target : 0, at, ns_off, o.desc != 0, aux) && comps_fit;
    };
    if (o.target < nk) { add(C_KEY, 0, ~0u); continue; }
    const AggOut& ao = pl->outs[o.target - nk];
    if (ao.dslot >= 0) { add(C_U64_AT, P.distinct[ao.dslot].dslot_off, ~0u); continue; }
    if (ao.kagg < 0) { add(C_COUNT, 0, ~0u); continue; }
    const AggDesc& d = sp.aggs[ao.kagg];
    // an all-NULL SUM/AVG/MIN/MAX orders first ascending, last descending
    // (compareAggregateValue, v2_shape_group_topn.zig:1361-1370)
    if (ao.op != TDBG_AGG_COUNT && d.g_ns_off != ~0u) add(C_AGG_NULL, 0, d.g_ns_off);
    switch (ao.op) {
      case TDBG_AGG_COUNT: add(C_U64_AT, d.g_off, ~0u); break;
      case TDBG_AGG_SUM:
        if (ao.is_float) add(C_F64_AT, d.g_off, d.g_ns_off);
        // a DECIMAL SUM known to fit in 64 bits orders as its low word, signed
        // (the same order as the i128's high word, then low word)
        else if (ao.out_type == TDBG_DECIMAL128 && sum_fits_i64(kaggs[ao.kagg].col)) add(C_I64_AT, d.g_off, d.g_ns_off);
        else if (ao.out_type == TDBG_DECIMAL128) { add(C_I128_HI, d.g_off, d.g_ns_off); add(C_I128_LO, d.g_off, d.g_ns_off); }
        else add(C_I64_AT, d.g_off, d.g_ns_off);  // BIGINT SUM orders by its wrapped value
        break;
      case TDBG_AGG_AVG:
        if (ao.is_float) add(C_AVG_F, d.g_off, d.g_ns_off);
        else add(C_AVG_I, d.g_off, d.g_ns_off, ao.scale);
        break;
      default: add(ao.is_float ? C_FORD_AT : C_I64_AT, d.g_off, d.g_ns_off); break;
    }
  }
  // Ties break by the silo's packed group key, ascending (rowOrderBefore,
  // v2_shape_group_topn.zig:1327-1335): its fields compare from the last key
  // part to the first, each as (validity, two's-complement bits).  Over key
  // parts whose values share one sign that is the order of the relative key
  // (the mixed-radix index, last part most significant, NULL = 0): a dense
  // table's slot order, which the sorts' final position tie-break keeps, or
  // the relative key stored in keyed records as one word.  Otherwise every
  // part's field is a word of its own.
  bool one_sign = true;
  for (uint32_t i = 0; i < sp.n_parts; i++) {
    const PartDesc& d = sp.parts[i];
    const __int128 lo = (__int128)d.base + (d.nullable ? 1 : 0);
    const __int128 hi = (__int128)d.base + (d.span ? (__int128)d.span : ((__int128)1 << 64)) - 1;
    if (lo <= hi && lo < 0 && hi >= 0) one_sign = false;
  }
  const bool keyed = sp.g_mode != G_DENSE;  // compacted radix leaves and the HBM hash table
  if (!op.n_comps && !(P.radix && !P.rx.dense_out)) {
    // no ORDER BY: nothing is sorted
  } else if (one_sign && keyed) {
    if (sp.g_mode == G_WIDE) comps_fit = add_comp(C_U64_AT, 0, sp.g_key_off + 8, ~0u, false, 0) && comps_fit;
    comps_fit = add_comp(C_U64_AT, 0, sp.g_key_off, ~0u, false, 0) && comps_fit;
  } else if (!one_sign) {
    for (int i = (int)sp.n_parts - 1; i >= 0; i--) {
      const uint32_t ty = pl->keys[i].type;
      const uint32_t bits = ty == TDBG_BOOLEAN || ty == TDBG_TINYINT ? 8u : ty == TDBG_SMALLINT ? 16u
                          : (ty == TDBG_INT || ty == TDBG_DATE) ? 32u : 64u;
      if (sp.parts[i].nullable) comps_fit = add_comp(C_PKEY_VALID, (uint32_t)i, 0, ~0u, false, 0) && comps_fit;
      comps_fit = add_comp(C_PKEY, (uint32_t)i, 0, ~0u, false, bits) && comps_fit;
    }
  }
  if (!comps_fit) return fail(TDBG_ERR_UNSUPPORTED, "ORDER BY needs more than %d order words", kMaxComps);
  // finalize writers; output buffer offsets are assigned per execution
  fp.n_outs = 0;
  for (uint32_t i = 0; i < sp.n_parts; i++) {
    OutDesc o{};
    o.kind = O_KEY; o.idx = i; o.bytes = (uint32_t)type_bytes(pl->keys[i].type);
    o.valid_off = sp.parts[i].nullable ? 0 : ~0ull; o.ns_off = ~0u; o.scale_div = 1.0;
    fp.outs[fp.n_outs++] = o;
  }
  for (auto& ao : pl->outs) {
    OutDesc o{};
    // SUM/AVG/MIN/MAX are NULL only for a group without a non-null input,
    // which a non-nullable input column cannot produce: no validity bytes
    o.ns_off = ~0u; o.valid_off = ao.nullable ? 0 : ~0ull; o.scale_div = 1.0;
    o.bytes = (uint32_t)type_bytes(ao.out_type);
    const AggDesc* d = ao.kagg >= 0 ? &sp.aggs[ao.kagg] : nullptr;
    if (d) { o.off = d->g_off; o.ns_off = d->g_ns_off; }
    switch (ao.op) {
      case TDBG_AGG_COUNT_STAR: o.kind = O_COUNT; o.valid_off = ~0ull; break;
      case TDBG_AGG_COUNT: o.kind = d ? O_NS : O_COUNT; o.valid_off = ~0ull; break;
      case TDBG_AGG_COUNT_DISTINCT:
        o.kind = O_NS;
        o.off = P.distinct[ao.dslot].dslot_off;
        o.valid_off = ~0ull;
        break;
      case TDBG_AGG_SUM:
        o.kind = ao.is_float ? O_SUM_F64 : (ao.out_type == TDBG_DECIMAL128 ? O_SUM_I128 : O_SUM_I64);
        break;
      case TDBG_AGG_AVG: {
        o.kind = ao.is_float ? O_AVG_F : O_AVG_I;
        double p10 = 1.0;
        for (uint32_t s = 0; s < ao.scale; s++) p10 *= 10.0;
        o.scale_div = p10;
        break;
      }
      default:
        o.kind = ao.is_float ? O_MINMAX_F : O_MINMAX_I;
        o.is_f32 = ao.in_type == TDBG_FLOAT;
        break;
    }
    fp.outs[fp.n_outs++] = o;
  }
  // HAVING in the leaf pass: every output's value must be in the leaf's LDS
  // table (a COUNT(DISTINCT) count is not: it is filled after compaction)
  if (!pl->having.empty() && P.radix && !P.rx.dense_out) {
    bool ok = true;
    HavingLeaf& L = P.hleaf;
    L.n = (uint32_t)pl->having.size();
    for (size_t k = 0; k < pl->having.size(); k++) memcpy(&L.nodes[k], &pl->having[k], sizeof(HavingNode));
    for (uint32_t c = 0; c < fp.n_outs && ok; c++) {
      const OutDesc& o = fp.outs[c];
      HavingSrc& hs = L.src[c];
      hs = HavingSrc{};
      hs.kind = o.kind;
      hs.part = o.idx;
      hs.scale_div = o.scale_div;
      if (o.kind == O_KEY || o.kind == O_COUNT) continue;
      if (c >= sp.n_parts && pl->outs[c - sp.n_parts].dslot >= 0) { ok = false; break; }
      hs.agg = ~0u;
      for (uint32_t k = 0; k < sp.n_aggs; k++)
        if (sp.aggs[k].g_off == o.off) hs.agg = k;
      if (hs.agg == ~0u) ok = false;
    }
    P.leaf_having = ok;
  }
  P.version = t->version;
  P.rows_hint = n_rows_scan;
  return TDBG_OK;
}

}  // namespace rt
}  // namespace tdbg

namespace {

int32_t sync_check(tdbg_plan* pl, const char* what) {
  hipError_t e = hipStreamSynchronize(pl->stream);
  if (e != hipSuccess) return hip_fail(e, what);
  return TDBG_OK;
}

}  // namespace

// ============================================================================
// C-ABI
// ============================================================================

extern "C" {

uint32_t tdbg_abi_version(void) { return TDBG_ABI_VERSION; }
const char* tdbg_last_error(void) { return g_last_error.c_str(); }

int32_t tdbg_context_create(int32_t device, uint64_t hbm_budget_bytes, tdbg_context** out) {
  if (!out) return fail(TDBG_ERR_INVALID, "null out");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(TDBG_ERR_DEVICE, "no HIP device: %s", hipGetErrorString(e));
  if (device < 0 || device >= n) return fail(TDBG_ERR_INVALID, "device %d of %d", device, n);
  HIP_TRY(hipSetDevice(device), "hipSetDevice");
  // the device's copy of the no-alloc scratch array, resolved once here
  if (int32_t rc = noalloc_setup(device)) return rc;
  auto* ctx = new tdbg_context();
  ctx->device = device;
  hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device);
  hipDeviceGetAttribute(&ctx->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  size_t free_b = 0, total_b = 0;
  hipMemGetInfo(&free_b, &total_b);
  ctx->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(total_b * 0.9);
  *out = ctx;
  return TDBG_OK;
}

void tdbg_context_destroy(tdbg_context* ctx) {
  if (!ctx) return;
  {
    std::unique_lock<std::shared_mutex> lk(ctx->mu);
    for (auto& kv : ctx->tables) free_table(ctx, kv.second.get());
    ctx->tables.clear();
    dev_free(ctx, ctx->d_stage, ctx->d_stage_bytes);
    ctx->d_stage = nullptr;
  }
  delete ctx;
}

uint64_t tdbg_context_bytes_in_use(const tdbg_context* ctx) { return ctx ? ctx->in_use.load() : 0; }

int32_t tdbg_table_reserve(tdbg_context* ctx, uint64_t table_uid, uint64_t rows) {
  if (!ctx) return fail(TDBG_ERR_INVALID, "null context");
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  std::unique_lock<std::shared_mutex> lk(ctx->mu);
  auto& tp = ctx->tables[table_uid];
  if (!tp) { tp.reset(new Table()); tp->uid = table_uid; tp->version = ++ctx->clock; }
  tp->reserve = std::max(tp->reserve, align_up(rows + kRowAlign * 64, kRowAlign));
  if (tp->cap_rows < tp->reserve && !tp->cols.empty()) return table_grow(ctx, tp.get(), tp->reserve);
  return TDBG_OK;
}

int32_t tdbg_table_drop(tdbg_context* ctx, uint64_t table_uid) {
  if (!ctx) return fail(TDBG_ERR_INVALID, "null context");
  std::unique_lock<std::shared_mutex> lk(ctx->mu);
  auto it = ctx->tables.find(table_uid);
  if (it == ctx->tables.end()) return TDBG_OK;
  free_table(ctx, it->second.get());
  ctx->tables.erase(it);
  return TDBG_OK;
}

}  // extern "C"

namespace {

// One block into its row group's rows of the column's arena: finds or
// creates both, lets `fill` write the values at the row group's offset
// (validity from `host_validity`, else all valid), then records the block's
// zone-map statistics.  Serialised by the context lock.
int32_t stage_block(tdbg_context* ctx, const tdbg_column_key* key, uint32_t type, const uint8_t* host_validity,
                    uint64_t n_rows, const tdbg_block_stats& stats,
                    const std::function<int32_t(uint8_t* dst, uint32_t elem)>& fill) {
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  std::unique_lock<std::shared_mutex> lk(ctx->mu);
  auto& tp = ctx->tables[key->table_uid];
  if (!tp) { tp.reset(new Table()); tp->uid = key->table_uid; tp->version = ++ctx->clock; }
  Table* t = tp.get();
  // row group entry
  RowGroupEntry* rg = nullptr;
  auto& bucket = t->rg_index[rg_hash(key->segment_id, key->row_group_idx)];
  for (size_t i : bucket)
    if (t->rgs[i].segment_id == key->segment_id && t->rgs[i].rg_idx == key->row_group_idx) rg = &t->rgs[i];
  if (rg && rg->n_rows != n_rows) return fail(TDBG_ERR_INVALID, "row group holds %u rows, block has %llu", rg->n_rows, (unsigned long long)n_rows);
  if (!rg) {
    const uint64_t off = align_up(t->next_row, kRowAlign);
    const uint64_t end = align_up(off + n_rows, kRowAlign);
    if (end > t->cap_rows) {
      int32_t rc = table_grow(ctx, t, end);
      if (rc) return rc;
    }
    RowGroupEntry e;
    e.segment_id = key->segment_id;
    e.rg_idx = key->row_group_idx;
    e.n_rows = (uint32_t)n_rows;
    e.row_off = off;
    t->rgs.push_back(e);
    bucket.push_back(t->rgs.size() - 1);
    rg = &t->rgs.back();
    t->next_row = off + n_rows;
  }
  // column arena
  auto ait = t->cols.find(key->column_idx);
  if (ait == t->cols.end()) {
    if (t->cap_rows == 0) {
      int32_t rc = table_grow(ctx, t, align_up(t->next_row, kRowAlign));
      if (rc) return rc;
    }
    ColumnArena a;
    a.type = type;
    a.elem = (uint32_t)type_bytes(type);
    void* p = nullptr;
    int32_t rc = dev_alloc(ctx, &p, arena_bytes(t->cap_rows, a.elem));
    if (rc) return rc;
    HIP_TRY(hipMemset(p, 0, arena_bytes(t->cap_rows, a.elem)), "arena clear");
    a.d_data = static_cast<uint8_t*>(p);
    ait = t->cols.emplace(key->column_idx, a).first;
  }
  ColumnArena& a = ait->second;
  if (a.type != type) return fail(TDBG_ERR_INVALID, "column %u staged as type %u, now %u", key->column_idx, a.type, type);
  if (host_validity && !a.d_valid) {
    void* p = nullptr;
    int32_t rc = dev_alloc(ctx, &p, valid_bytes(t->cap_rows));
    if (rc) return rc;
    HIP_TRY(hipMemset(p, 0xff, valid_bytes(t->cap_rows)), "validity init");
    a.d_valid = static_cast<uint32_t*>(p);
  }
  int32_t rc = fill(a.d_data + rg->row_off * a.elem, a.elem);
  if (rc) return rc;
  if (a.d_valid) {
    uint8_t* vb = reinterpret_cast<uint8_t*>(a.d_valid) + rg->row_off / 8;
    if (host_validity) HIP_TRY(hipMemcpy(vb, host_validity, (n_rows + 7) / 8, hipMemcpyHostToDevice), "stage validity");
    else HIP_TRY(hipMemset(vb, 0xff, (n_rows + 7) / 8), "stage validity");
  }
  BlockInfo& bi = rg->blocks[key->column_idx];
  bi.staged = true;
  bi.stats = stats;
  bi.has_stats = bi.stats.has_minmax != 0;
  if (bi.has_stats) {
    if (!a.has_stats) {
      a.min_i = bi.stats.min_i; a.max_i = bi.stats.max_i;
      a.min_f = bi.stats.min_f; a.max_f = bi.stats.max_f;
      a.has_stats = true;
    } else {
      a.min_i = std::min(a.min_i, bi.stats.min_i); a.max_i = std::max(a.max_i, bi.stats.max_i);
      a.min_f = std::min(a.min_f, bi.stats.min_f); a.max_f = std::max(a.max_f, bi.stats.max_f);
    }
  }
  t->version = ++ctx->clock;  // unique across tables and drops: plans cache by it
  return TDBG_OK;
}

// The device staging buffer of encoded bodies (caller holds ctx->mu).
int32_t stage_scratch(tdbg_context* ctx, uint64_t bytes, uint8_t** out) {
  if (ctx->d_stage_bytes < bytes) {
    dev_free(ctx, ctx->d_stage, ctx->d_stage_bytes);
    ctx->d_stage = nullptr;
    ctx->d_stage_bytes = 0;
    void* p = nullptr;
    const uint64_t want = std::max<uint64_t>(align_up(bytes, 1 << 20), 1 << 20);
    int32_t rc = dev_alloc(ctx, &p, want);
    if (rc) return rc;
    ctx->d_stage = p;
    ctx->d_stage_bytes = want;
  }
  *out = static_cast<uint8_t*>(ctx->d_stage);
  return TDBG_OK;
}

uint64_t rd_le(const uint8_t* p, uint32_t bytes) {
  uint64_t v = 0;
  for (uint32_t b = 0; b < bytes; b++) v |= (uint64_t)p[b] << (8 * b);
  return v;
}

}  // namespace

extern "C" {

int32_t tdbg_column_stage(tdbg_context* ctx, const tdbg_column_key* key, uint32_t type, const void* host_values,
                          const uint8_t* host_validity, uint64_t n_rows, const tdbg_block_stats* stats) {
  if (!ctx || !key) return fail(TDBG_ERR_INVALID, "null argument");
  if (!is_int_type(type) && !is_float_type(type)) return fail(TDBG_ERR_UNSUPPORTED, "column type %u", type);
  if (n_rows == 0 || n_rows > 0xffffffffull) return fail(TDBG_ERR_INVALID, "row count %llu", (unsigned long long)n_rows);
  if (!host_values) return fail(TDBG_ERR_INVALID, "null values");
  tdbg_block_stats st{};
  if (stats && stats->has_minmax) st = *stats;
  else host_stats(type, host_values, host_validity, n_rows, &st);
  return stage_block(ctx, key, type, host_validity, n_rows, st, [&](uint8_t* dst, uint32_t elem) -> int32_t {
    HIP_TRY(hipMemcpy(dst, host_values, n_rows * elem, hipMemcpyHostToDevice), "stage values");
    return TDBG_OK;
  });
}

int32_t tdbg_column_stage_encoded(tdbg_context* ctx, const tdbg_column_key* key, uint32_t type, uint32_t encoding,
                                  const void* payload, uint64_t payload_bytes, uint64_t n_rows, uint32_t has_nulls,
                                  const tdbg_block_stats* stats) {
  if (!ctx || !key || !payload) return fail(TDBG_ERR_INVALID, "null argument");
  if (!is_int_type(type) && !is_float_type(type)) return fail(TDBG_ERR_UNSUPPORTED, "column type %u", type);
  if (n_rows == 0 || n_rows > 0xffffffffull) return fail(TDBG_ERR_INVALID, "row count %llu", (unsigned long long)n_rows);
  const uint8_t* p = static_cast<const uint8_t*>(payload);
  const uint64_t bm = has_nulls ? (n_rows + 7) / 8 : 0;  // column.bitmapBytes
  if (payload_bytes < bm) return fail(TDBG_ERR_INVALID, "payload of %llu bytes shorter than its bitmap", (unsigned long long)payload_bytes);
  const uint8_t* valid = has_nulls ? p : nullptr;
  const uint8_t* body = p + bm;
  const uint64_t body_bytes = payload_bytes - bm;
  const uint32_t elem = (uint32_t)type_bytes(type);
  if (encoding == TDBG_ENC_RAW) {
    if (body_bytes < n_rows * elem) return fail(TDBG_ERR_INVALID, "raw body of %llu bytes for %llu rows", (unsigned long long)body_bytes, (unsigned long long)n_rows);
    return tdbg_column_stage(ctx, key, type, body, valid, n_rows, stats);
  }
  if (encoding != TDBG_ENC_FOR && encoding != TDBG_ENC_RLE)
    return fail(TDBG_ERR_UNSUPPORTED, "encoding %u (dict/fsst carry strings, not on this path)", encoding);
  if (is_float_type(type)) return fail(TDBG_ERR_UNSUPPORTED, "FOR/RLE blocks hold integer-family columns only");
  // host view of the values, for zone-map statistics when the caller has none
  std::vector<uint8_t> host;
  auto need_host = [&]() { return !(stats && stats->has_minmax); };
  if (encoding == TDBG_ENC_FOR) {
    // [base i128 LE][width u8][3 pad][codes: n_rows x width LE] (forBlockOf)
    if (body_bytes < 20) return fail(TDBG_ERR_INVALID, "FOR header truncated");
    const int64_t base = (int64_t)rd_le(body, 8);
    const uint64_t base_hi = rd_le(body + 8, 8);
    if (base_hi != (base < 0 ? ~0ull : 0ull)) return fail(TDBG_ERR_INVALID, "FOR base beyond 64 bits");
    const uint32_t width = body[16];
    if (width != 1 && width != 2 && width != 4 && width != 8) return fail(TDBG_ERR_INVALID, "FOR width %u", width);
    if (body_bytes < 20 + n_rows * width) return fail(TDBG_ERR_INVALID, "FOR codes truncated");
    const uint8_t* codes = body + 20;
    tdbg_block_stats st{};
    if (!need_host()) st = *stats;
    else {
      host.resize(n_rows * elem);
      for (uint64_t r = 0; r < n_rows; r++) {
        const uint64_t v = (uint64_t)base + rd_le(codes + r * width, width);
        memcpy(host.data() + r * elem, &v, elem);  // little-endian: the low `elem` bytes, wrapped
      }
      host_stats(type, host.data(), valid, n_rows, &st);
    }
    return stage_block(ctx, key, type, valid, n_rows, st, [&](uint8_t* dst, uint32_t el) -> int32_t {
      uint8_t* d = nullptr;
      int32_t rc = stage_scratch(ctx, n_rows * width, &d);
      if (rc) return rc;
      HIP_TRY(hipMemcpy(d, codes, n_rows * width, hipMemcpyHostToDevice), "stage FOR codes");
      HIP_TRY(launch_for_expand(d, width, base, el, n_rows, dst, nullptr), "FOR expand");
      HIP_TRY(hipStreamSynchronize(nullptr), "FOR expand");
      return TDBG_OK;
    });
  }
  // RLE: [n_runs u32][value_width u8][3 pad][values: n_runs x width][lengths: n_runs x u32] (rleBlockOf)
  if (body_bytes < 8) return fail(TDBG_ERR_INVALID, "RLE header truncated");
  const uint32_t n_runs = (uint32_t)rd_le(body, 4);
  const uint32_t width = body[4];
  if (width != elem) return fail(TDBG_ERR_INVALID, "RLE value width %u for a %u-byte column", width, elem);
  if (body_bytes < 8 + (uint64_t)n_runs * (width + 4)) return fail(TDBG_ERR_INVALID, "RLE runs truncated");
  const uint8_t* vals = body + 8;
  const uint8_t* lens = vals + (uint64_t)n_runs * width;
  // run starts, clamped to the block (fillRuns): runs past the last row drop
  std::vector<uint32_t> starts;
  uint64_t pos = 0;
  uint32_t used = 0;
  for (; used < n_runs && pos < n_rows; used++) {
    starts.push_back((uint32_t)pos);
    pos += std::min<uint64_t>(rd_le(lens + (uint64_t)used * 4, 4), n_rows - pos);
  }
  const uint32_t covered = (uint32_t)pos;
  tdbg_block_stats st{};
  if (!need_host()) st = *stats;
  else {
    host.assign(n_rows * elem, 0);
    for (uint32_t r = 0; r < used; r++) {
      const uint64_t end = r + 1 < used ? starts[r + 1] : covered;
      for (uint64_t i = starts[r]; i < end; i++) memcpy(host.data() + i * elem, vals + (uint64_t)r * width, elem);
    }
    host_stats(type, host.data(), valid, n_rows, &st);
  }
  return stage_block(ctx, key, type, valid, n_rows, st, [&](uint8_t* dst, uint32_t el) -> int32_t {
    const uint64_t vals_b = align_up((uint64_t)used * width, 256);
    uint8_t* d = nullptr;
    int32_t rc = stage_scratch(ctx, vals_b + (uint64_t)used * 4 + 256, &d);
    if (rc) return rc;
    if (used) {
      HIP_TRY(hipMemcpy(d, vals, (uint64_t)used * width, hipMemcpyHostToDevice), "stage RLE values");
      HIP_TRY(hipMemcpy(d + vals_b, starts.data(), (uint64_t)used * 4, hipMemcpyHostToDevice), "stage RLE runs");
    }
    HIP_TRY(launch_rle_expand(d, reinterpret_cast<const uint32_t*>(d + vals_b), used, covered, el, n_rows, dst, nullptr),
            "RLE expand");
    HIP_TRY(hipStreamSynchronize(nullptr), "RLE expand");
    return TDBG_OK;
  });
}

int32_t tdbg_column_evict(tdbg_context* ctx, const tdbg_column_key* key) {
  if (!ctx || !key) return fail(TDBG_ERR_INVALID, "null argument");
  std::unique_lock<std::shared_mutex> lk(ctx->mu);
  auto it = ctx->tables.find(key->table_uid);
  if (it == ctx->tables.end()) return TDBG_OK;
  Table* t = it->second.get();
  auto b = t->rg_index.find(rg_hash(key->segment_id, key->row_group_idx));
  if (b == t->rg_index.end()) return TDBG_OK;
  for (size_t i : b->second) {
    RowGroupEntry& rg = t->rgs[i];
    if (rg.segment_id != key->segment_id || rg.rg_idx != key->row_group_idx) continue;
    auto bi = rg.blocks.find(key->column_idx);
    if (bi != rg.blocks.end()) bi->second.staged = false;
  }
  t->version = ++ctx->clock;  // unique across tables and drops: plans cache by it
  return TDBG_OK;
}

int32_t tdbg_plan_create(tdbg_context* ctx, const tdbg_plan_desc* d, tdbg_plan** out) {
  if (!ctx || !d || !out) return fail(TDBG_ERR_INVALID, "null argument");
  *out = nullptr;
  if ((d->n_keys && !d->keys) || (d->n_aggs && !d->aggs) || (d->n_preds && !d->preds) || (d->n_order && !d->order) ||
      (d->n_having && !d->having) || (d->n_where && !d->where))
    return fail(TDBG_ERR_INVALID, "null array in plan");
  std::unique_ptr<tdbg_plan> pl(new tdbg_plan());
  pl->ctx = ctx;
  pl->table_uid = d->table_uid;
  pl->keys.assign(d->keys, d->keys + d->n_keys);
  pl->aggs.assign(d->aggs, d->aggs + d->n_aggs);
  pl->preds.assign(d->preds, d->preds + d->n_preds);
  pl->order.assign(d->order, d->order + d->n_order);
  if (d->n_having) pl->having.assign(d->having, d->having + d->n_having);
  if (d->n_where) pl->where.assign(d->where, d->where + d->n_where);
  pl->limit = d->limit;
  pl->offset = d->offset;
  pl->est_groups = d->est_groups;
  int32_t rc = validate(pl.get());
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  // device resources; a failure part way releases what was made
  auto setup = [&]() -> int32_t {
    HIP_TRY(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking), "stream");
    for (auto& e : pl->ev) HIP_TRY(hipEventCreate(&e), "event");
    if (!pl->having.empty()) {
      const uint64_t hb = pl->having.size() * sizeof(tdbg_having_node);
      int32_t r = ensure(pl.get(), pl->having_nodes, hb);
      if (r) return r;
      HIP_TRY(hipMemcpy(pl->having_nodes.p, pl->having.data(), hb, hipMemcpyHostToDevice), "HAVING nodes");
    }
    return TDBG_OK;
  };
  if ((rc = setup())) {
    tdbg_plan_destroy(pl.release());
    return rc;
  }
  *out = pl.release();
  return TDBG_OK;
}

void tdbg_plan_destroy(tdbg_plan* pl) {
  if (!pl) return;
  (void)hipSetDevice(pl->ctx->device);
  if (pl->stream) (void)hipStreamSynchronize(pl->stream);
  for (auto* b : {&pl->state, &pl->tiles, &pl->err, &pl->gids, &pl->counts, &pl->sel, &pl->scratch, &pl->out, &pl->tmpl, &pl->nout,
                 &pl->rx_misc, &pl->rx_rows1, &pl->rx_rows2, &pl->rx_leaf, &pl->rx_words, &pl->having_nodes,
                 &pl->having_gids, &pl->dist_set, &pl->dist_counts, &pl->dist_entries, &pl->having_leaf})
    release(pl, *b);
  for (auto& e : pl->ev) if (e) (void)hipEventDestroy(e);
  if (pl->h_in) (void)hipHostFree(pl->h_in);
  if (pl->stream) (void)hipStreamDestroy(pl->stream);
  delete pl;
}

int32_t tdbg_cancel(tdbg_plan* pl) {
  if (!pl) return fail(TDBG_ERR_INVALID, "null plan");
  pl->cancel_gen.fetch_add(1);
  return TDBG_OK;
}

int32_t tdbg_plan_last_stats(const tdbg_plan* pl, tdbg_exec_stats* out) {
  if (!pl || !out) return fail(TDBG_ERR_INVALID, "null argument");
  *out = pl->last;
  return TDBG_OK;
}

}  // extern "C"

int32_t tdbg::rt::make_result(tdbg_plan* pl, uint64_t n_out, const uint8_t* host, const std::vector<uint64_t>& data_off,
                              const std::vector<uint64_t>& valid_off, std::shared_ptr<void> keep, tdbg_result** out) {
  auto* r = new tdbg_result();
  r->keep = keep;
  r->n_rows = n_out;
  const size_t nk = pl->keys.size();
  const size_t nc = nk + pl->outs.size();
  r->cols.resize(nc);
  r->data.resize(nc);
  r->valid.resize(nc);
  for (size_t c = 0; c < nc; c++) {
    tdbg_column_view& rc = r->cols[c];
    uint32_t type, scale = 0, nullable = 0;
    if (c < nk) {
      type = pl->keys[c].type;
      nullable = pl->prep.sp.parts[c].nullable;
    } else {
      const AggOut& o = pl->outs[c - nk];
      type = o.out_type;
      if (o.out_type == TDBG_DECIMAL128 || o.out_type == TDBG_DECIMAL64) scale = o.scale;
      nullable = o.op != TDBG_AGG_COUNT_STAR && o.op != TDBG_AGG_COUNT && o.op != TDBG_AGG_COUNT_DISTINCT;
    }
    const uint64_t bytes = n_out * (uint64_t)type_bytes(type);
    if (!keep) {
      r->data[c].resize(bytes + 16);
      if (n_out) memcpy(r->data[c].data(), host + data_off[c], bytes);
    }
    bool any_null = false;
    if (valid_off[c] != ~0ull && n_out) {
      std::vector<uint8_t>& bm = r->valid[c];
      bm.assign((n_out + 7) / 8, 0);
      const uint8_t* vb = host + valid_off[c];
      for (uint64_t i = 0; i < n_out; i++) {
        if (vb[i]) bm[i >> 3] |= (uint8_t)(1u << (i & 7));
        else any_null = true;
      }
      if (!any_null) bm.clear();
    }
    rc.type = type;
    rc.scale = scale;
    rc.nullable = nullable;
    rc.data = keep ? host + data_off[c] : r->data[c].data();
    rc.validity = any_null ? r->valid[c].data() : nullptr;
    rc.n_rows = n_out;
  }
  *out = r;
  return TDBG_OK;
}

// COUNT(DISTINCT dd.col) of the plan's groups (distinct.hip): a set of
// (group, value) composites with at least twice as many slots as rows
// scanned, a counter per relative key, then each compacted group's count.
int32_t tdbg::rt::run_distinct(tdbg_plan* pl, Table* t, const ScanSet& S, const Prepared::Distinct& dd,
                               const uint32_t* gids, const uint64_t* n_groups, uint64_t max_groups) {
  Prepared& P = pl->prep;
  const ScanParams& sp = P.sp;
  hipStream_t s = pl->stream;
  auto it = t->cols.find(dd.col);
  if (it == t->cols.end()) return fail(TDBG_ERR_NOT_STAGED, "column %u not staged", dd.col);
  const ColumnArena& a = it->second;
  unsigned __int128 dom = 1;
  for (uint32_t i = 0; i < sp.n_parts; i++) dom *= sp.parts[i].span;
  // a scan's records are addressable by key (dense index or the HBM hash
  // table); radix leaves write them where their rows started, so radix
  // plans count per relative key and gather afterwards
  const bool to_records = !P.radix;
  if (!to_records && dom > kDistinctDomainMax)
    return fail(TDBG_ERR_UNSUPPORTED, "COUNT(DISTINCT) over a key domain of more than 2^31 groups");
  DistinctParams D{};
  D.to_records = to_records ? 1u : 0u;
  D.x.data = a.d_data;
  D.x.valid = a.d_valid;
  D.x.kind = kind_of(a.type);
  D.is_float = is_float_type(a.type) ? 1u : 0u;
  D.base = a.has_stats ? a.min_i : 0;
  D.span = a.has_stats ? (uint64_t)a.max_i - (uint64_t)a.min_i + 1 : 1;
  const unsigned __int128 groups = to_records ? (unsigned __int128)sp.g_cap : dom;  // composite's group index
  // a composite key * span + rel when it fits 63 bits, else (a float, a
  // wide span) entries of (group, value bits) behind the slots
  D.wide = (D.is_float || D.span == 0 || groups * D.span >= ((unsigned __int128)1 << 63)) ? 1u : 0u;
  D.dslot_off = dd.dslot_off;
  const uint64_t pairs = D.wide ? std::max<uint64_t>(S.rows, 1)
                                : std::min<uint64_t>(std::max<uint64_t>(S.rows, 1), (uint64_t)(groups * D.span));
  uint64_t slots = 1024;
  while (slots < 2 * pairs) slots <<= 1;
  // a composite range small enough for a bit array (up to 2^33 bits, and
  // no larger than the hash set) takes one bit per pair: no probing, and a
  // footprint the device's last-level cache can hold
  const unsigned __int128 bits = groups * D.span;
  D.bitmap = (!D.wide && bits <= ((unsigned __int128)1 << 33) && (uint64_t)(bits / 8) <= slots * 8) ? 1u : 0u;
  const uint64_t words = D.bitmap ? (uint64_t)((bits + 31) / 32) : 0;
  int32_t rc;
  if ((rc = ensure(pl, pl->dist_set, D.bitmap ? words * 4 + 256 : slots * 8))) return rc;
  if (D.wide) {  // one entry per arena row a tile may hold
    if ((rc = ensure(pl, pl->dist_entries, std::max<uint64_t>(t->next_row, 1) * 16 + 256))) return rc;
    D.entries = static_cast<uint64_t*>(pl->dist_entries.p);
  }
  D.set = static_cast<unsigned long long*>(pl->dist_set.p);
  D.set_mask = slots - 1;
  if (D.bitmap) HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(D.set), words, 0u, s), "distinct bit array");
  else HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(D.set), slots * 2, ~0u, s), "distinct set");
  if (!to_records) {
    if ((rc = ensure(pl, pl->dist_counts, (uint64_t)dom * 4 + 256))) return rc;
    D.counts = static_cast<uint32_t*>(pl->dist_counts.p);
    HIP_TRY(launch_fill_u32(D.counts, (uint64_t)dom, 0, s), "distinct counts");
  }
  // LDS counters (small domains) are flushed once per workgroup: fewer, longer-lived workgroups
  const uint32_t per_cu = (uint64_t)dom <= kDistinctLdsKeys ? 4u : 8u;
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(sp.n_tiles, (uint32_t)std::max(pl->ctx->n_cu, 8) * per_cu));
  // (the LDS counters need the domain below 2^32; larger domains never take them)
  HIP_TRY(launch_distinct_insert(sp, D, (uint64_t)std::min<unsigned __int128>(dom, ~0u), grid, s), "distinct insert");
  if (!to_records)
    HIP_TRY(launch_distinct_gather(P.fp, D, gids, n_groups, max_groups, s), "distinct counts to groups");
  return TDBG_OK;
}

namespace {

// tdbg_execute_noalloc: for one execution, the plan's device buffers are
// set aside and ensure() carves fresh ones from the scratch region; the
// plan's own buffers come back afterwards, untouched.
struct ArenaScope {
  tdbg_plan* pl;
  tdbg_plan::Buf saved[20];
  static constexpr tdbg_plan::Buf tdbg_plan::*kBufs[20] = {
      &tdbg_plan::state, &tdbg_plan::tiles, &tdbg_plan::err, &tdbg_plan::gids, &tdbg_plan::counts,
      &tdbg_plan::sel, &tdbg_plan::scratch, &tdbg_plan::out, &tdbg_plan::tmpl, &tdbg_plan::nout,
      &tdbg_plan::rx_misc, &tdbg_plan::rx_rows1, &tdbg_plan::rx_rows2, &tdbg_plan::rx_leaf, &tdbg_plan::rx_words,
      &tdbg_plan::having_gids, &tdbg_plan::dist_set, &tdbg_plan::dist_counts, &tdbg_plan::having_leaf,
      &tdbg_plan::dist_entries};  // every buffer but having_nodes, which the plan uploaded at creation
  ArenaScope(tdbg_plan* p, uint8_t* base, uint64_t cap) : pl(p) {
    for (int i = 0; i < 20; i++) {
      saved[i] = pl->*kBufs[i];
      pl->*kBufs[i] = tdbg_plan::Buf{};
    }
    pl->arena = tdbg_plan::Arena{base, cap, 0};
    pl->arena_on = true;
  }
  ~ArenaScope() {
    for (int i = 0; i < 20; i++) pl->*kBufs[i] = saved[i];
    pl->arena_on = false;
    pl->arena = tdbg_plan::Arena{};
  }
};

int32_t run_plan(tdbg_plan* pl, const tdbg_rg_ref* rgs, uint32_t n_rgs, bool views, tdbg_result** out,
                 const NoAllocOut* na, uint64_t gen);

// run_plan's request to run again after re-planning (never returned to a
// caller: the loop ends because each retry doubles the leaves, up to 2^20)
constexpr int32_t kRetryMoreLeaves = -1000;

}  // namespace

int32_t tdbg::rt::execute_plan(tdbg_plan* pl, const tdbg_rg_ref* rgs, uint32_t n_rgs, bool views, tdbg_result** out,
                               uint64_t cancel_gen) {
  if (!pl || !out) return fail(TDBG_ERR_INVALID, "null argument");
  *out = nullptr;
  int32_t rc;
  while ((rc = run_plan(pl, rgs, n_rgs, views, out, nullptr, cancel_gen)) == kRetryMoreLeaves) {}
  return rc;
}

int32_t tdbg::rt::execute_plan_noalloc(tdbg_plan* pl, const tdbg_rg_ref* rgs, uint32_t n_rgs, const NoAllocOut& na) {
  if (!pl || !na.cols || !na.n_rows || (!na.out && na.out_capacity)) return fail(TDBG_ERR_INVALID, "null argument");
  *na.n_rows = 0;
  if (na.n_cols != pl->keys.size() + pl->outs.size())
    return fail(TDBG_ERR_INVALID, "%u column views for a plan of %zu output columns", na.n_cols,
                pl->keys.size() + pl->outs.size());
  const uint64_t gen = pl->cancel_gen.load();
  int32_t rc;
  while ((rc = run_plan(pl, rgs, n_rgs, false, nullptr, &na, gen)) == kRetryMoreLeaves) {}
  return rc;
}

namespace {

// The result's column views into host memory holding, for column c, n_out
// values at data_off[c] and (valid_off[c] != ~0) one validity byte per row,
// which is rewritten in place as the views' bitmap (bit r of byte r / 8).
void views_in_place(tdbg_plan* pl, uint64_t n_out, uint8_t* host, const std::vector<uint64_t>& data_off,
                    const std::vector<uint64_t>& valid_off, tdbg_column_view* cols) {
  const size_t nk = pl->keys.size();
  const size_t nc = nk + pl->outs.size();
  for (size_t c = 0; c < nc; c++) {
    tdbg_column_view& v = cols[c];
    v = tdbg_column_view{};
    if (c < nk) {
      v.type = pl->keys[c].type;
      v.nullable = pl->prep.sp.parts[c].nullable;
    } else {
      const AggOut& o = pl->outs[c - nk];
      v.type = o.out_type;
      if (o.out_type == TDBG_DECIMAL128 || o.out_type == TDBG_DECIMAL64) v.scale = o.scale;
      v.nullable = o.op != TDBG_AGG_COUNT_STAR && o.op != TDBG_AGG_COUNT && o.op != TDBG_AGG_COUNT_DISTINCT;
    }
    v.n_rows = n_out;
    v.data = n_out ? host + data_off[c] : nullptr;
    if (valid_off[c] == ~0ull || !n_out) continue;
    uint8_t* vb = host + valid_off[c];
    bool any_null = false;
    for (uint64_t j = 0; j < (n_out + 7) / 8; j++) {  // byte j is written after bytes 8j.. are read
      uint8_t m = 0;
      for (uint64_t i = 8 * j; i < std::min<uint64_t>(n_out, 8 * j + 8); i++) {
        if (vb[i]) m |= (uint8_t)(1u << (i & 7));
        else any_null = true;
      }
      vb[j] = m;
    }
    v.validity = any_null ? vb : nullptr;
  }
}

int32_t run_plan(tdbg_plan* pl, const tdbg_rg_ref* rgs, uint32_t n_rgs, bool views, tdbg_result** out,
                 const NoAllocOut* na, uint64_t gen) {
  tdbg_context* ctx = pl->ctx;
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  // lock order everywhere: the context's table lock, then the plan's
  std::shared_lock<std::shared_mutex> lk(ctx->mu);
  std::lock_guard<std::mutex> plk(pl->mu);
  std::unique_ptr<ArenaScope> arena;
  if (na) arena.reset(new ArenaScope(pl, na->scratch, na->scratch_bytes));
  if (cancelled(pl, gen)) return fail(TDBG_ERR_CANCELLED, "cancelled before it started");
  pl->last = tdbg_exec_stats{};
  auto tit = ctx->tables.find(pl->table_uid);
  if (tit == ctx->tables.end()) return fail(TDBG_ERR_NOT_STAGED, "table %llu has no staged columns", (unsigned long long)pl->table_uid);
  Table* t = tit->second.get();
  // the row groups' tiles, reused while the table and the row-group list
  // are unchanged (resolving thousands of row groups per query is host time)
  int32_t rc;
  tdbg_plan::ScanCache& C = pl->scan_cache;
  const bool hit = C.valid && C.table == t && C.version == t->version && C.all == (rgs == nullptr) &&
                   (rgs == nullptr || (C.rgs.size() == n_rgs &&
                                       (n_rgs == 0 || memcmp(C.rgs.data(), rgs, n_rgs * sizeof(tdbg_rg_ref)) == 0)));
  if (!hit) {
    C.valid = false;
    if ((rc = collect_scan(pl, t, rgs, n_rgs, C.S))) return rc;
    C.table = t;
    C.version = t->version;
    C.all = rgs == nullptr;
    if (rgs) C.rgs.assign(rgs, rgs + n_rgs);
    else C.rgs.clear();
    C.valid = true;
  }
  const ScanSet& S = C.S;
  const std::vector<Tile>& tiles = S.tiles;
  const uint64_t rows = S.rows;
  // The plan specialised to this scan is reused while the scan set is (and
  // nothing else - a batch's hints - prepared the plan since, and the
  // planner's switches read the same).
  std::string env_now;
  for (const char* v : {"THINDB_GPU_RADIX", "THINDB_GPU_RADIX_LEAF_BITS", "THINDB_GPU_RADIX_BITS1", "THINDB_GPU_RADIX_HASH",
                        "THINDB_GPU_RADIX_HASH_TAG_BITS", "THINDB_GPU_RADIX_HASH_EXACT", "THINDB_GPU_RADIX_HASH_WGS",
                        "THINDB_GPU_RADIX_HASH_FILL"}) {
    const char* e = getenv(v);
    env_now += e ? e : "-";
    env_now += '|';
  }
  if (!(hit && pl->exec_prep_stamp == pl->prep_stamp && pl->exec_prep_env == env_now)) {
    if ((rc = prepare(pl, t, S, PlanHint{}))) return rc;
    pl->exec_prep_stamp = pl->prep_stamp;
    pl->exec_prep_env = env_now;
  }
  Prepared& P = pl->prep;
  ScanParams& sp = P.sp;
  // offsets of the output buffer
  const size_t nk = pl->keys.size();
  const size_t nc = nk + pl->outs.size();
  std::vector<uint64_t> data_off(nc, 0), valid_off(nc, ~0ull);
  const uint64_t need_lim = pl->limit == 0 ? ~0ull : pl->offset + pl->limit;
  if (tiles.empty()) {
    pl->last.rows_scanned = 0;
    if (na) {
      views_in_place(pl, 0, nullptr, data_off, valid_off, na->cols);
      return TDBG_OK;
    }
    return make_result(pl, 0, nullptr, data_off, valid_off, nullptr, out);
  }
  hipStream_t s = pl->stream;
  // buffers
  const uint64_t state_bytes = sp.g_cap * sp.g_rec;
  if ((rc = ensure(pl, pl->state, state_bytes))) return rc;
  if ((rc = ensure(pl, pl->tiles, tiles.size() * sizeof(Tile)))) return rc;
  if ((rc = ensure(pl, pl->err, 64))) return rc;
  if ((rc = ensure(pl, pl->tmpl, P.tmpl.size() * 8))) return rc;
  if ((rc = ensure(pl, pl->nout, 64))) return rc;
  sp.g_state = static_cast<uint8_t*>(pl->state.p);
  sp.tiles = static_cast<const Tile*>(pl->tiles.p);
  sp.n_tiles = (uint32_t)tiles.size();
  sp.err = static_cast<uint32_t*>(pl->err.p);
  P.op.g_state = sp.g_state;
  P.fp.g_state = sp.g_state;
  const uint64_t scanned_bytes = S.bytes;

  // tiles and the record template go through the plan's pinned staging
  // buffer, so the copies are asynchronous (the previous execution's copies
  // completed at its final synchronisation)
  const uint64_t tiles_b = tiles.size() * sizeof(Tile), tmpl_b = P.tmpl.size() * 8;
  const uint64_t hl_at = align_up(align_up(tiles_b, 256) + tmpl_b, 256);
  const uint64_t hl_b = P.leaf_having ? sizeof(HavingLeaf) : 0;
  // (no-alloc: pageable host memory that outlives the copies, which end at
  // the final synchronisation)
  std::vector<uint8_t> h_pageable;
  uint8_t* h_in = nullptr;
  if (na) {
    h_pageable.resize(hl_at + hl_b);
    h_in = h_pageable.data();
  } else {
    if ((rc = ensure_pinned(pl->h_in, pl->h_in_cap, hl_at + hl_b))) return rc;
    h_in = pl->h_in;
  }
  if (hl_b && (rc = ensure(pl, pl->having_leaf, hl_b))) return rc;
  // the device copies stay valid while the preparation and the buffers do
  // (a no-alloc execution carves fresh buffers: it always uploads)
  const bool uploaded = !na && pl->up_stamp == pl->prep_stamp && pl->up_tiles == pl->tiles.p &&
                        pl->up_tmpl == pl->tmpl.p && pl->up_hleaf == (hl_b ? pl->having_leaf.p : nullptr);
  if (!uploaded) {
    memcpy(h_in, tiles.data(), tiles_b);
    memcpy(h_in + align_up(tiles_b, 256), P.tmpl.data(), tmpl_b);
    HIP_TRY(hipMemcpyAsync(pl->tiles.p, h_in, tiles_b, hipMemcpyHostToDevice, s), "tiles");
    HIP_TRY(hipMemcpyAsync(pl->tmpl.p, h_in + align_up(tiles_b, 256), tmpl_b, hipMemcpyHostToDevice, s), "template");
    if (hl_b) {
      memcpy(h_in + hl_at, &P.hleaf, hl_b);
      HIP_TRY(hipMemcpyAsync(pl->having_leaf.p, h_in + hl_at, hl_b, hipMemcpyHostToDevice, s), "HAVING");
    }
    if (!na) {
      pl->up_stamp = pl->prep_stamp;
      pl->up_tiles = pl->tiles.p;
      pl->up_tmpl = pl->tmpl.p;
      pl->up_hleaf = hl_b ? pl->having_leaf.p : nullptr;
    }
  }
  bool specialized = false;
  const uint64_t n_rec = sp.g_cap;
  if ((rc = ensure(pl, pl->gids, n_rec * 4 + 256))) return rc;
  if (P.radix) {
    // partition buffers: rows (twice when two levels), level-1 histogram,
    // bucket starts and cursors, the emitted-group cursor, leaf ranges
    RadixParams& R = P.rx;
    const uint64_t nb1 = 1ull << R.bits1;
    const uint64_t row_bytes = rows * R.row.stride + 256;
    if ((rc = ensure(pl, pl->rx_rows1, row_bytes + (R.bits2 ? rows * 4 + 256 : 0)))) return rc;
    if (R.bits2 && (rc = ensure(pl, pl->rx_rows2, row_bytes))) return rc;
    // leaf_groups: n_leaves counts, a total, then one sum per 256 leaves;
    // then the groups' count histogram
    const uint64_t lg_words = (R.n_leaves + 1 + R.n_leaves / 256 + 2) / 2 + 1;
    // then the group counters of per-leaf top-k
    const uint64_t total_at = nb1 + (nb1 + 1) + nb1 + lg_words + kRxCountBins / 2;
    const uint64_t misc_words = total_at + 8ull * kRxTotalSlots;
    if ((rc = ensure(pl, pl->rx_misc, misc_words * 8 + 256))) return rc;
    // level 2: about four workgroups per CU over all buckets
    R.chunks2 = R.bits2 ? (uint32_t)std::min<uint64_t>(64, std::max<uint64_t>(1, 4ull * ctx->n_cu / nb1)) : 0;
    const uint64_t hist2_b = R.bits2 ? nb1 * R.chunks2 * (1ull << R.bits2) * 4 : 0;
    if (R.bits2 && (rc = ensure(pl, pl->rx_leaf, 2ull * R.n_leaves * 8 + 256 + hist2_b))) return rc;
    uint64_t* misc = static_cast<uint64_t*>(pl->rx_misc.p);
    R.bucket_start = misc + nb1;
    R.bucket_fill = R.bucket_start + nb1 + 1;
    R.leaf_groups = reinterpret_cast<uint32_t*>(R.bucket_fill + nb1);
    R.rows1 = static_cast<uint8_t*>(pl->rx_rows1.p);
    // two levels: the level-1 routes after the rows (the level-2 count reads them alone)
    R.digits1 = R.bits2 ? reinterpret_cast<uint32_t*>(R.rows1 + align_up(row_bytes, 256)) : nullptr;
    R.rows2 = R.bits2 ? static_cast<uint8_t*>(pl->rx_rows2.p) : R.rows1;
    R.leaf_start = R.bits2 ? static_cast<uint64_t*>(pl->rx_leaf.p) : nullptr;
    R.hist2 = R.bits2 ? reinterpret_cast<uint32_t*>(R.leaf_start + 2ull * R.n_leaves + 32) : nullptr;
    // ORDER BY COUNT(*) ... LIMIT over compacted leaves (the reference's
    // native COUNT top-N, v2_shape_group_topn.zig:846-870): the leaves also
    // write each group's count and a histogram of them, which fixes the cut
    // without another pass over the groups
    R.having = P.leaf_having ? static_cast<const HavingLeaf*>(pl->having_leaf.p) : nullptr;
    R.counts_out = R.count_hist = nullptr;
    R.leaf_topk = R.topk_desc = 0;
    R.group_total = nullptr;
    // ORDER BY COUNT(*) with a small LIMIT, ties broken by the relative key
    // (one word: the key parts share a sign), which within a leaf is slot
    // order: each leaf writes only its best need_lim groups by (count, slot),
    // so the groups outside every leaf's top never reach HBM
    // (THINDB_GPU_LEAF_TOPK=0: every group, for tests that compare the two)
    const char* lt_env = getenv("THINDB_GPU_LEAF_TOPK");
    const bool count_then_key = P.op.n_comps == 2 && P.op.comps[0].src == C_COUNT && P.op.comps[1].src == C_U64_AT &&
                                P.op.comps[1].off == sp.g_key_off && !P.op.comps[1].desc;
    if (!R.dense_out && !R.hashed && count_then_key && need_lim > 0 && need_lim <= kLeafTopMax &&
        pl->having.empty() && P.distinct.empty() && sp.lds_cap <= 64u * 512u && !(lt_env && lt_env[0] == '0')) {
      R.leaf_topk = (uint32_t)need_lim;
      R.topk_desc = P.op.comps[0].desc != 0;
      R.group_total = reinterpret_cast<unsigned long long*>(misc + total_at);
    }
    if (!R.dense_out && P.op.n_comps > 0 && P.op.comps[0].src == C_COUNT && need_lim <= kSortCandidatesMax &&
        (rows > 65536 || R.leaf_topk) && (pl->having.empty() || P.leaf_having)) {
      if ((rc = ensure(pl, pl->rx_words, rows * 4 + 256))) return rc;
      R.counts_out = static_cast<uint32_t*>(pl->rx_words.p);
      R.count_hist = reinterpret_cast<uint32_t*>(R.bucket_fill + nb1 + lg_words);
    }
    auto* hist = reinterpret_cast<unsigned long long*>(misc);
    HIP_TRY(hipEventRecord(pl->ev[0], s), "event");
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(misc), misc_words * 2, 0, s), "init histogram");
    HIP_TRY(launch_fill_u32(sp.err, 16, 0, s), "init flags");
    HIP_TRY(hipEventRecord(pl->ev[1], s), "event");
    HIP_TRY(launch_radix(sp, R, std::min<uint32_t>(2 * P.grid, sp.n_tiles), hist, s, RX_HIST), "radix histogram");
    HIP_TRY(launch_radix(sp, R, 1, hist, s, RX_OFFSETS), "radix offsets");
    HIP_TRY(launch_radix(sp, R, P.grid, hist, s, RX_SCATTER1), "radix scatter");
    if (R.bits2) HIP_TRY(launch_radix(sp, R, 1, hist, s, RX_SCATTER2), "radix scatter 2");
    HIP_TRY(launch_radix(sp, R, P.agg_grid, hist, s, RX_AGG), "radix aggregate");
    HIP_TRY(hipEventRecord(pl->ev[2], s), "event");
    if (!R.dense_out)
      HIP_TRY(launch_radix_gids(R, static_cast<uint32_t*>(pl->gids.p), static_cast<uint64_t*>(pl->nout.p), s),
              "group positions");
  } else {
    HIP_TRY(hipEventRecord(pl->ev[0], s), "event");
    HIP_TRY(launch_fill_records(sp.g_state, sp.g_cap, sp.g_rec, static_cast<const uint64_t*>(pl->tmpl.p), s, sp.err,
                                16), "init records and flags");
    HIP_TRY(hipEventRecord(pl->ev[1], s), "event");
    if (launch_scan_jit(sp, P.grid, P.block, s, &specialized, na != nullptr) != hipSuccess || !specialized) {
      (void)hipGetLastError();
      HIP_TRY(launch_scan_agg(sp, P.grid, P.block, s), "scan_agg launch");
    }
    HIP_TRY(hipEventRecord(pl->ev[2], s), "event");
  }
  if (cancelled(pl, gen)) { (void)hipStreamSynchronize(s); return fail(TDBG_ERR_CANCELLED, "cancelled"); }
  // A dense table of at most kFastRec records without HAVING or
  // COUNT(DISTINCT), ordered with a small OFFSET + LIMIT (or small and
  // unordered): one workgroup orders, slices and writes the result, and one
  // copy brings it back with the group count and the error flags, so the
  // query synchronises once.  THINDB_GPU_FAST_FINISH=0 keeps the general path.
  {
    const char* ff_env = getenv("THINDB_GPU_FAST_FINISH");
    const uint64_t rows_cap = std::min<uint64_t>(need_lim, n_rec);
    const bool fast = !na && !(ff_env && ff_env[0] == '0') && !P.radix && sp.g_mode == G_DENSE &&
                      P.distinct.empty() && pl->having.empty() && n_rec <= kFastRec && pl->offset < rows_cap &&
                      P.op.n_comps <= kFastWords &&
                      (P.op.n_comps ? need_lim <= kFastNeed : (need_lim <= kFastNeed || n_rec <= kFastUnordered));
    if (fast) {
      const uint64_t cap_out = rows_cap - pl->offset;
      uint64_t ob = 0;
      for (size_t c = 0; c < nc; c++) {
        OutDesc& o = P.fp.outs[c];
        ob = align_up(ob, 16);
        data_off[c] = o.data_off = ob;
        ob += cap_out * o.bytes;
        if (o.valid_off != ~0ull) {
          ob = align_up(ob, 16);
          valid_off[c] = o.valid_off = ob;
          ob += cap_out;
        }
      }
      const uint64_t head = 256;  // FastHeader, then the columns
      // the kernel writes the header and the rows straight into the pinned
      // result buffer (through its device address): no copy after it
      if (!pl->h_out.p || pl->h_out.p.use_count() > 1 || pl->h_out.cap < head + ob + 16) {
        pl->h_out = PinnedBuf{};
        if ((rc = pinned_acquire(ctx->pinned, head + ob + 16, pl->h_out))) return rc;
      }
      uint8_t* h = pl->h_out.p.get();
      void* dh = nullptr;
      HIP_TRY(hipHostGetDevicePointer(&dh, h, 0), "pinned result address");
      uint8_t* d_out = static_cast<uint8_t*>(dh);
      HIP_TRY(launch_dense_topn(P.op, P.fp, (uint32_t)n_rec, (uint32_t)pl->offset, (uint32_t)rows_cap, sp.err,
                                d_out + head, reinterpret_cast<FastHeader*>(d_out), s), "dense top-N");
      HIP_TRY(hipEventRecord(pl->ev[3], s), "event");
      if ((rc = sync_check(pl, "scan"))) return rc;
      const FastHeader hd = *reinterpret_cast<const FastHeader*>(h);
      if (hd.err & ERR_KEY_RANGE) return fail(TDBG_ERR_INVALID, "a key value lies outside the staged zone-map statistics");
      if (hd.err & ERR_HASH_FULL) return fail(TDBG_ERR_OOM, "group table full: est_groups too small");
      if (cancelled(pl, gen)) return fail(TDBG_ERR_CANCELLED, "cancelled");
      if (!hd.fallback) {
        if (hd.n_out > cap_out) return fail(TDBG_ERR_DEVICE, "dense top-N wrote %llu rows of %llu",
                                            (unsigned long long)hd.n_out, (unsigned long long)cap_out);
        float ms_total = 0, ms_scan = 0;
        hipEventElapsedTime(&ms_total, pl->ev[0], pl->ev[3]);
        hipEventElapsedTime(&ms_scan, pl->ev[1], pl->ev[2]);
        pl->last.total_ms = ms_total;
        pl->last.scan_ms = ms_scan;
        pl->last.rows_scanned = rows;
        pl->last.bytes_scanned = scanned_bytes;
        pl->last.n_groups = hd.n_groups;
        pl->last.lds_mode = sp.lds_mode;
        pl->last.global_mode = sp.g_mode;
        pl->last.specialized = specialized ? 1u : 0u;
        pl->last.lds_parts = sp.lds_mode == LDS_DIRECT ? sp.lds_parts : 1u;
        return make_result(pl, hd.n_out, h + head, data_off, valid_off,
                           views ? std::shared_ptr<void>(pl->h_out.p, h) : nullptr, out);
      }
      // more than kFastCand groups tie at the cut: the general path below
    }
  }
  // compaction (compacted radix leaves listed their groups' positions above)
  if (!P.radix || P.rx.dense_out) {
    if ((rc = ensure(pl, pl->counts, ((n_rec + 4095) / 4096 + 1) * 4))) return rc;
    HIP_TRY(launch_compact(P.op, n_rec, static_cast<uint32_t*>(pl->counts.p), static_cast<uint32_t*>(pl->gids.p),
                           static_cast<uint64_t*>(pl->nout.p), s), "compact");
  }
  uint32_t* gids = static_cast<uint32_t*>(pl->gids.p);
  uint64_t* nout = static_cast<uint64_t*>(pl->nout.p);
  // COUNT(DISTINCT): one insertion pass per input column, then each group's
  // count into its record
  for (const Prepared::Distinct& dd : P.distinct) {
    if ((rc = run_distinct(pl, t, S, dd, gids, nout, n_rec))) return rc;
  }
  // HAVING: the passing groups' positions, in order, to having_gids and their
  // number to nout[4]
  const bool having = !pl->having.empty() && !P.leaf_having;  // else the leaf pass applied it
  if (having) {
    // survivors' gids, then the pass masks and per-workgroup counts
    const uint64_t out_b = align_up(n_rec * 4 + 256, 256);
    if ((rc = ensure(pl, pl->having_gids, out_b + having_scratch_bytes(n_rec)))) return rc;
    uint8_t* hb = static_cast<uint8_t*>(pl->having_gids.p);
    HIP_TRY(launch_having(P.fp, static_cast<const HavingNode*>(pl->having_nodes.p), (uint32_t)pl->having.size(), gids,
                          nout, n_rec, hb + out_b, reinterpret_cast<uint32_t*>(hb), nout + 4, s), "having");
  }
  uint64_t n_groups = 0, n_kept = 0;
  uint32_t err_bits = 0;
  uint32_t count_hist[kRxCountBins];
  const bool count_topn = P.radix && P.rx.counts_out;
  // per-leaf top-k: the leaves wrote candidates only (n_groups below, their
  // counts and count histogram as above); every group was counted
  unsigned long long group_total[8 * kRxTotalSlots];
  const bool leaf_topk = P.radix && P.rx.leaf_topk;
  HIP_TRY(hipMemcpyAsync(&n_groups, pl->nout.p, 8, hipMemcpyDeviceToHost, s), "group count");
  if (leaf_topk)
    HIP_TRY(hipMemcpyAsync(group_total, P.rx.group_total, sizeof group_total, hipMemcpyDeviceToHost, s),
            "group total");
  if (having) HIP_TRY(hipMemcpyAsync(&n_kept, nout + 4, 8, hipMemcpyDeviceToHost, s), "HAVING count");
  HIP_TRY(hipMemcpyAsync(&err_bits, sp.err, 4, hipMemcpyDeviceToHost, s), "error flags");
  if (count_topn)
    HIP_TRY(hipMemcpyAsync(count_hist, P.rx.count_hist, sizeof count_hist, hipMemcpyDeviceToHost, s), "count histogram");
  if ((rc = sync_check(pl, "scan"))) return rc;
  if (err_bits & ERR_KEY_RANGE) return fail(TDBG_ERR_INVALID, "a key value lies outside the staged zone-map statistics");
  if ((err_bits & ERR_HASH_FULL) && P.radix && P.rx.hashed && P.rx.n_leaves < (1u << 20)) {
    // a leaf held more groups than its table: partition into twice the leaves
    pl->rx_hash_boost++;
    pl->exec_prep_stamp = ~0ull;
    return kRetryMoreLeaves;
  }
  if (err_bits & ERR_HASH_FULL) return fail(TDBG_ERR_OOM, "group table full: est_groups too small");
  if (cancelled(pl, gen)) return fail(TDBG_ERR_CANCELLED, "cancelled");
  if (P.radix && !P.rx.dense_out && n_groups > n_rec)
    return fail(TDBG_ERR_DEVICE, "radix pass emitted %llu groups for %llu records", (unsigned long long)n_groups,
                (unsigned long long)n_rec);
  if (having) {
    if (n_kept > n_groups)
      return fail(TDBG_ERR_DEVICE, "HAVING kept %llu of %llu groups", (unsigned long long)n_kept,
                  (unsigned long long)n_groups);
    n_groups = n_kept;
    gids = static_cast<uint32_t*>(pl->having_gids.p);
  }
  // ORDER BY / OFFSET / LIMIT
  const uint64_t start = std::min<uint64_t>(pl->offset, n_groups);
  const uint64_t end = std::min<uint64_t>(need_lim, n_groups);
  const uint64_t n_out = end - start;
  const uint32_t* sel = gids + start;
  if (P.op.n_comps > 0 && n_groups > 0 && end > 0) {
    if ((rc = ensure(pl, pl->sel, end * 4 + 256))) return rc;
    // COUNT(*) first: the cut is the count at which the histogram, walked
    // from the best end, reaches `end` groups; counts in the last bin are
    // not resolved, so a cut there takes the general path
    bool done = false;
    if (count_topn) {
      const bool desc = P.op.comps[0].desc != 0;
      uint64_t cum = 0;
      uint32_t cut = 0;
      for (uint32_t k = 1; k < kRxCountBins; k++) {
        const uint32_t b = desc ? kRxCountBins - k : k;
        cum += count_hist[b];
        if (cum >= end) {
          cut = b;
          break;
        }
      }
      if (cut != 0 && cut != kRxCountBins - 1 && cum <= kSortCandidatesMax) {
        const uint32_t m = (uint32_t)cum;
        const size_t cb = align_up(sort_candidates_scratch_bytes(m, P.op.n_comps), 256);
        if ((rc = ensure(pl, pl->scratch, cb + (uint64_t)m * 4 + 256))) return rc;
        uint32_t* idx = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(pl->scratch.p) + cb);
        HIP_TRY(launch_radix_count_take(P.rx, cut, desc, idx, m, reinterpret_cast<uint32_t*>(pl->nout.p) + 2, s),
                "count cut");
        HIP_TRY(launch_sort_candidates(P.op, gids, n_groups, idx, m, end, pl->scratch.p,
                                       static_cast<uint32_t*>(pl->sel.p), s), "order");
        done = true;
      }
    }
    if (!done) {
      if ((rc = ensure(pl, pl->scratch, sort_scratch_bytes(n_groups, P.op.n_comps)))) return rc;
      HIP_TRY(launch_sort(P.op, gids, n_groups, end, pl->scratch.p, pl->scratch.bytes,
                          static_cast<uint32_t*>(pl->sel.p), s), "order");
    }
    sel = static_cast<const uint32_t*>(pl->sel.p) + start;
  }
  // finalize into one output buffer
  uint64_t ob = 0;
  for (size_t c = 0; c < nc; c++) {
    FinalParams& fp = P.fp;
    OutDesc& o = fp.outs[c];
    ob = align_up(ob, 16);
    data_off[c] = ob;
    o.data_off = ob;
    ob += n_out * o.bytes;
    if (o.valid_off != ~0ull) {
      ob = align_up(ob, 16);
      valid_off[c] = ob;
      o.valid_off = ob;
      ob += n_out;
    }
  }
  if ((rc = ensure(pl, pl->out, ob + 256))) return rc;
  HIP_TRY(launch_finalize(P.fp, sel, n_out, static_cast<uint8_t*>(pl->out.p), s), "finalize");
  HIP_TRY(hipEventRecord(pl->ev[3], s), "event");
  if (na) {  // straight into the caller's host buffer
    if (ob > na->out_capacity) {
      (void)hipStreamSynchronize(s);
      return fail(TDBG_ERR_INVALID, "result of %llu rows needs %llu output bytes, %llu given",
                  (unsigned long long)n_out, (unsigned long long)ob, (unsigned long long)na->out_capacity);
    }
    if (n_out) HIP_TRY(hipMemcpyAsync(na->out, pl->out.p, ob, hipMemcpyDeviceToHost, s), "result copy");
    if ((rc = sync_check(pl, "finalize"))) return rc;
  } else {
    // the result lands in pinned memory its columns then view in place; a
    // buffer a live earlier result still views is left to it
    if (!pl->h_out.p || pl->h_out.p.use_count() > 1 || pl->h_out.cap < ob + 16) {
      pl->h_out = PinnedBuf{};  // back to the pool unless a result still views it
      if ((rc = pinned_acquire(ctx->pinned, ob + 16, pl->h_out))) return rc;
    }
    if (n_out) HIP_TRY(hipMemcpyAsync(pl->h_out.p.get(), pl->out.p, ob, hipMemcpyDeviceToHost, s), "result copy");
    if ((rc = sync_check(pl, "finalize"))) return rc;
  }
  float ms_total = 0, ms_scan = 0;
  hipEventElapsedTime(&ms_total, pl->ev[0], pl->ev[3]);
  hipEventElapsedTime(&ms_scan, pl->ev[1], pl->ev[2]);
  pl->last.total_ms = ms_total;
  pl->last.scan_ms = ms_scan;
  pl->last.rows_scanned = rows;
  pl->last.bytes_scanned = scanned_bytes;
  uint64_t n_all = n_groups;
  if (leaf_topk) {
    n_all = 0;
    for (uint32_t i = 0; i < kRxTotalSlots; i++) n_all += group_total[8 * i];
  }
  pl->last.n_groups = n_all;
  pl->last.lds_mode = sp.lds_mode;
  pl->last.global_mode = P.radix ? (P.rx.hashed ? 4u : 3u) : sp.g_mode;
  pl->last.specialized = specialized ? 1u : 0u;
  pl->last.lds_parts = sp.lds_mode == LDS_DIRECT ? sp.lds_parts : 1u;
  if (na) {
    views_in_place(pl, n_out, static_cast<uint8_t*>(na->out), data_off, valid_off, na->cols);
    *na->n_rows = n_out;
    return TDBG_OK;
  }
  return make_result(pl, n_out, pl->h_out.p.get(), data_off, valid_off,
                     views ? std::shared_ptr<void>(pl->h_out.p, pl->h_out.p.get()) : nullptr, out);
}

}  // namespace

extern "C" {

int32_t tdbg_execute(tdbg_plan* pl, const tdbg_rg_ref* rgs, uint32_t n_rgs, tdbg_result** out) {
  // a tdbg_cancel from here on (while this call waits for the plan or runs) stops it
  return execute_plan(pl, rgs, n_rgs, true, out, pl ? pl->cancel_gen.load() : 0);
}

uint64_t tdbg_result_num_rows(const tdbg_result* r) { return r ? r->n_rows : 0; }
uint32_t tdbg_result_num_columns(const tdbg_result* r) { return r ? (uint32_t)r->cols.size() : 0; }
int32_t tdbg_result_column(const tdbg_result* r, uint32_t i, tdbg_column_view* out) {
  if (!r || !out || i >= r->cols.size()) return fail(TDBG_ERR_INVALID, "column %u", i);
  *out = r->cols[i];
  return TDBG_OK;
}
void tdbg_result_free(tdbg_result* r) { delete r; }

}  // extern "C"
