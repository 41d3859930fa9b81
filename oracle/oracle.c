/*
 * oracle.c — CPU restatement of thinDB's grouped aggregate (scan -> filter
 * -> GROUP BY -> ORDER BY/LIMIT), the checker for the GPU engine.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg; never by the product.
 *
 * What it restates (thinDB source tree, read as text):
 *   - the request gate and agg planning: src/exec/v2_lowcard_group.zig:234-373
 *     (declines: >8 keys, >16 aggs, non-integer keys, temporal SUM/AVG,
 *     decimal128 inputs) and the silo's <=128-bit key packing,
 *     src/exec/v2_shape_group_topn.zig:2051-2094;
 *   - the silo's high-cardinality arrangement, for the CPU baseline only
 *     (oracle_group_by_radix, end of file): routeHashKey bucket routing, the
 *     staged bucket rows and the deferred parts each folded into a table of
 *     its own, src/exec/silo_group_core.zig:5218-5235, :9228-9400,
 *     :10799-10960 - the same result as oracle_group_by, tested equal;
 *   - key packing: packIntPart, v2_lowcard_group.zig:886-948
 *     (`(bits(v) & mask(width)) << offset`, widths from keyTypeBits :184-192);
 *   - the worker model: tiles of TILE_RGS=16 row groups claimed from a shared
 *     cursor (rg_tiles.zig:1-30; v2_lowcard_group.zig:839-854), scanned in
 *     sub-batches of autoScanBatch rows (exec.zig:77-85);
 *   - runs of equal adjacent keys and one gid per run: groupRows/runEnds,
 *     v2_lowcard_group.zig:1054-1175, with the direct gid array of
 *     resolveRunsDirect :1095-1118 / directRangeOf :509-530;
 *   - the fold kernels: foldCountCol/foldSumInt/foldSumIntDense/
 *     foldSumFloat/foldExtremeInt/foldExtremeFloat, :1206-1347 (run
 *     accumulators, i128 integer slots, f64 per-run sums, first value seeds
 *     MIN/MAX);
 *   - the worker merge: mergeScalars, :1938-1999 (i128 add, f64 add, min/max
 *     with the `had` rule);
 *   - ORDER BY: the silo's result order, rowOrderBefore/compareOrderKey/
 *     compareAggregateValue, src/exec/v2_shape_group_topn.zig:1327-1395
 *     (NULL keys and all-NULL aggregates first ascending, BIGINT SUM by its
 *     wrapped value, AVG by the emitted f64(sum)/n/10^scale, floats by
 *     types.floatOrder, src/types.zig:579-587; ties broken by the packed
 *     group key, so the order is total).  The low-cardinality core sorts by
 *     the same items (SortCtx.lessThan/compareBySpec, v2_lowcard_group.zig:
 *     2198-2271, which never sees a NULL aggregate: it declines nullable
 *     inputs, :319-330) but breaks ties by gid, an order its parallel merge
 *     leaves unspecified; the packed key is one order it admits;
 *   - WHERE trees (AND / OR / NOT / IS [NOT] NULL over `col op literal`
 *     leaves): evaluatePredicate, predicate.zig:1727-1790;
 *   - COUNT(DISTINCT): the (group, value) composites of rows with a non-NULL
 *     integer value, deduplicated and counted per group (foldDistinctValues /
 *     compositeOf, v2_lowcard_group.zig:1359-1418, :1497-1499);
 *   - HAVING: havingPasses/havingOutputNum/numOrder,
 *     src/exec/v2_shape_group_topn.zig:2385-2535 (groups filtered before
 *     ORDER BY; a NULL operand fails its leaf; NOT is two-valued; integers
 *     compare as i128, anything with a double in f64 under floatOrder; the
 *     emitted value of each output: integerAggValue :1448-1452, avgValue
 *     :1438-1445, aggIsNull :1411-1423);
 *   - OFFSET/LIMIT: emitSorted :2069-2087 (limit 0 = all);
 *   - emit: emitGroup/appendInt/appendFloat :2089-2311 and the output types
 *     of src/exec/aggregate.zig:3700-3735 (COUNT -> BIGINT, SUM int -> BIGINT
 *     wrapped by @truncate, SUM decimal -> DECIMAL128, SUM float -> DOUBLE,
 *     AVG -> DOUBLE = f64(sum)/f64(n)/10^scale, MIN/MAX -> input type,
 *     NULL when the group has no non-null input).
 *   - hashing: mix64/hashU128 and the tiered key hash of IntKeyMemsetTable,
 *     src/exec/group_table.zig:45-56, :737-790; capacityFor :1711-1717.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

#define MAX_KEYS 8
#define MAX_AGGS 16
#define TILE_RGS 16
#define EMPTY_GID 0xffffffffu
#define DIRECT_MAX_SPAN (1u << 18)
#define DIRECT_SPAN_PER_GROUP 64u

static __thread char g_err[512];
static void set_err(const char* msg) { snprintf(g_err, sizeof g_err, "%s", msg); }
const char* oracle_last_error(void) { return g_err; }

/* ---- types ------------------------------------------------------------ */

static int type_bytes(uint32_t t) {
  switch (t) {
    case TDBG_BOOLEAN: case TDBG_TINYINT: return 1;
    case TDBG_SMALLINT: return 2;
    case TDBG_INT: case TDBG_DATE: case TDBG_FLOAT: return 4;
    case TDBG_BIGINT: case TDBG_DATETIME: case TDBG_DECIMAL64: case TDBG_DOUBLE: return 8;
    case TDBG_DECIMAL128: return 16;
    default: return 0;
  }
}
/* keyTypeBits, v2_lowcard_group.zig:184-192 */
static int key_type_bits(uint32_t t) {
  switch (t) {
    case TDBG_BOOLEAN: case TDBG_TINYINT: return 8;
    case TDBG_SMALLINT: return 16;
    case TDBG_INT: case TDBG_DATE: return 32;
    case TDBG_BIGINT: case TDBG_DATETIME: case TDBG_DECIMAL64: return 64;
    default: return 0;
  }
}
static int is_float_type(uint32_t t) { return t == TDBG_FLOAT || t == TDBG_DOUBLE; }

static inline int64_t load_i(const or_column* c, uint64_t r) {
  switch (c->type) {
    case TDBG_BOOLEAN: return ((const uint8_t*)c->data)[r];
    case TDBG_TINYINT: return ((const int8_t*)c->data)[r];
    case TDBG_SMALLINT: return ((const int16_t*)c->data)[r];
    case TDBG_INT: case TDBG_DATE: return ((const int32_t*)c->data)[r];
    default: return ((const int64_t*)c->data)[r];
  }
}
static inline double load_f(const or_column* c, uint64_t r) {
  if (c->type == TDBG_FLOAT) return (double)((const float*)c->data)[r];
  return ((const double*)c->data)[r];
}
static inline int is_valid(const or_column* c, uint64_t r) {
  if (!c->validity) return 1;
  return (c->validity[r >> 3] >> (r & 7)) & 1;
}

/* ---- hashing (group_table.zig:45-56) ------------------------------------ */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* ---- compiled plan -------------------------------------------------------- */

typedef struct {
  uint32_t col;
  uint32_t type;
  uint32_t offset, width;
  int nullable;       /* the key column has a validity bitmap */
  uint32_t valid_bit; /* set in the packed key for a non-NULL value */
} key_part;

enum { OP_COUNT_STAR, OP_COUNT_COL, OP_SUM, OP_AVG, OP_MIN, OP_MAX, OP_COUNT_DISTINCT };

typedef struct {
  int op;
  int col; /* -1 for COUNT(*) */
  int is_float;
  uint32_t in_type;
  uint32_t out_type;
  uint32_t scale;
  int nullable;
} agg_plan;

typedef struct {
  const or_table* t;
  const tdbg_plan_desc* d;
  key_part parts[MAX_KEYS];
  uint32_t n_parts;
  uint32_t key_bits;
  agg_plan aggs[MAX_AGGS];
  uint32_t n_aggs;
  uint32_t batch_rows;
  /* direct gid array for a lone key part (directRangeOf) */
  int direct;
  uint64_t direct_base, direct_mask;
  uint32_t direct_span;
  uint32_t dop;  /* workers (emit's top-k selection splits its groups over them) */
} cplan;

/* ---- per-worker state ----------------------------------------------------- */

typedef struct {
  /* open-addressing table, capacityFor/0.75 load (group_table.zig) */
  u128* tkeys;
  uint32_t* tgids;
  uint64_t cap, len;
  /* per-gid state, stride n_aggs (v2_lowcard_group.zig WState) */
  u128* keys;
  uint64_t* counts;
  i128* slots; /* integer states; floats live in fslots */
  double* fslots;
  double* csum; /* Neumaier: running sum + compensation per float agg */
  double* ccomp;
  uint64_t* ns;
  uint64_t n_groups, gcap;
  uint32_t* direct_gids;
  /* scratch */
  uint64_t* sel;
  u128* bkeys;
  uint32_t* ends;
  uint32_t* run_gids;
  int err;
} wstate;

static int ws_init(wstate* w, const cplan* p, uint64_t est) {
  memset(w, 0, sizeof *w);
  uint64_t need = est * 4 / 3 + 1, cap = 16;
  while (cap < need) cap *= 2;
  w->cap = cap;
  w->tkeys = malloc(cap * sizeof(u128));
  w->tgids = malloc(cap * sizeof(uint32_t));
  if (!w->tkeys || !w->tgids) return -1;
  for (uint64_t i = 0; i < cap; i++) w->tgids[i] = EMPTY_GID;
  if (p->direct) {
    w->direct_gids = malloc((size_t)p->direct_span * sizeof(uint32_t));
    if (!w->direct_gids) return -1;
    for (uint32_t i = 0; i < p->direct_span; i++) w->direct_gids[i] = EMPTY_GID;
  }
  size_t b = p->batch_rows + 8;
  w->sel = malloc(b * sizeof(uint64_t));
  w->bkeys = malloc(b * sizeof(u128));
  w->ends = malloc(b * sizeof(uint32_t));
  w->run_gids = malloc(b * sizeof(uint32_t));
  if (!w->sel || !w->bkeys || !w->ends || !w->run_gids) return -1;
  return 0;
}

static void ws_free(wstate* w) {
  free(w->tkeys); free(w->tgids); free(w->keys); free(w->counts); free(w->slots);
  free(w->fslots); free(w->csum); free(w->ccomp); free(w->ns); free(w->direct_gids);
  free(w->sel); free(w->bkeys); free(w->ends); free(w->run_gids);
}

static uint64_t hash_key(const cplan* p, u128 key) {
  uint64_t lo = (uint64_t)key, hi = (uint64_t)(key >> 64);
  if (p->key_bits <= 32) return mix64((uint64_t)(uint32_t)lo);
  if (p->key_bits <= 96) return mix64(lo ^ mix64((uint64_t)(uint32_t)hi));
  return mix64(lo ^ mix64(hi));
}

static int ws_grow_groups(wstate* w, uint32_t n_aggs) {
  uint64_t ng = w->gcap ? w->gcap * 2 : 64;
  w->keys = realloc(w->keys, ng * sizeof(u128));
  w->counts = realloc(w->counts, ng * sizeof(uint64_t));
  w->slots = realloc(w->slots, ng * n_aggs * sizeof(i128) + 16);
  w->fslots = realloc(w->fslots, ng * n_aggs * sizeof(double) + 8);
  w->csum = realloc(w->csum, ng * n_aggs * sizeof(double) + 8);
  w->ccomp = realloc(w->ccomp, ng * n_aggs * sizeof(double) + 8);
  w->ns = realloc(w->ns, ng * n_aggs * sizeof(uint64_t) + 8);
  if (!w->keys || !w->counts || !w->slots || !w->fslots || !w->csum || !w->ccomp || !w->ns) return -1;
  w->gcap = ng;
  return 0;
}

static int ws_table_grow(wstate* w, const cplan* p) {
  uint64_t nc = w->cap * 2;
  u128* nk = malloc(nc * sizeof(u128));
  uint32_t* ng = malloc(nc * sizeof(uint32_t));
  if (!nk || !ng) { free(nk); free(ng); return -1; }
  for (uint64_t i = 0; i < nc; i++) ng[i] = EMPTY_GID;
  for (uint64_t i = 0; i < w->cap; i++) {
    if (w->tgids[i] == EMPTY_GID) continue;
    uint64_t h = hash_key(p, w->tkeys[i]) & (nc - 1);
    while (ng[h] != EMPTY_GID) h = (h + 1) & (nc - 1);
    nk[h] = w->tkeys[i];
    ng[h] = w->tgids[i];
  }
  free(w->tkeys); free(w->tgids);
  w->tkeys = nk; w->tgids = ng; w->cap = nc;
  return 0;
}

/* getOrPut + addGroup (v2_lowcard_group.zig:1182-1198): a new gid starts
 * with zeroed accumulators. */
static uint32_t ws_get_or_add(wstate* w, const cplan* p, u128 key) {
  uint64_t mask = w->cap - 1;
  uint64_t h = hash_key(p, key) & mask;
  while (w->tgids[h] != EMPTY_GID) {
    if (w->tkeys[h] == key) return w->tgids[h];
    h = (h + 1) & mask;
  }
  if (w->n_groups == w->gcap && ws_grow_groups(w, p->n_aggs)) { w->err = 1; return 0; }
  uint32_t gid = (uint32_t)w->n_groups++;
  w->keys[gid] = key;
  w->counts[gid] = 0;
  for (uint32_t i = 0; i < p->n_aggs; i++) {
    size_t o = (size_t)gid * p->n_aggs + i;
    w->slots[o] = 0; w->fslots[o] = 0; w->csum[o] = 0; w->ccomp[o] = 0; w->ns[o] = 0;
  }
  if ((w->len + 1) * 4 > w->cap * 3) {
    if (ws_table_grow(w, p)) { w->err = 1; return 0; }
    mask = w->cap - 1;
    h = hash_key(p, key) & mask;
    while (w->tgids[h] != EMPTY_GID) h = (h + 1) & mask;
  }
  w->tkeys[h] = key;
  w->tgids[h] = gid;
  w->len++;
  return gid;
}

static inline void neumaier_add(double* s, double* c, double x) {
  double t = *s + x;
  if (fabs(*s) >= fabs(x)) *c += (*s - t) + x; else *c += (x - t) + *s;
  *s = t;
}

/* ---- filter (predicate.zig leaves, AND-ed) ---------------------------------- */
static inline int pred_pass(const tdbg_predicate* pr, const or_column* c, uint64_t r) {
  if (!is_valid(c, r)) return 0; /* NULL fails every comparison */
  if (is_float_type(pr->type)) {
    double v = load_f(c, r), x = pr->f64;
    switch (pr->op) {
      case TDBG_EQ: return v == x;
      case TDBG_NEQ: return v != x;
      case TDBG_LT: return v < x;
      case TDBG_LTE: return v <= x;
      case TDBG_GT: return v > x;
      default: return v >= x;
    }
  }
  int64_t v = load_i(c, r), x = pr->i64;
  switch (pr->op) {
    case TDBG_EQ: return v == x;
    case TDBG_NEQ: return v != x;
    case TDBG_LT: return v < x;
    case TDBG_LTE: return v <= x;
    case TDBG_GT: return v > x;
    default: return v >= x;
  }
}

/* The packed key of row r: each part's low `width` bits at its offset. */
static u128 pack_key(const cplan* p, uint64_t r) {
  u128 key = 0;
  for (uint32_t k = 0; k < p->n_parts; k++) {
    const key_part* kp = &p->parts[k];
    const or_column* c = &p->t->cols[kp->col];
    if (kp->nullable && !is_valid(c, r)) continue; /* a NULL packs value 0, its validity bit clear */
    uint64_t raw = (uint64_t)load_i(c, r);
    uint64_t m = kp->width >= 64 ? ~0ull : ((1ull << kp->width) - 1);
    key |= (u128)(raw & m) << kp->offset;
    if (kp->nullable) key |= (u128)1 << kp->valid_bit;
  }
  return key;
}

/* The WHERE tree's subtree rooted at postfix node `at` over row r
 * (evaluatePredicate, predicate.zig:1727-1790: a NULL never matches a
 * comparison, NOT inverts, AND/OR over `arity` children); the subtree's
 * first node index goes to *first. */
static int where_eval(const cplan* p, uint64_t r, int32_t at, int32_t* first) {
  const tdbg_filter_node* f = &p->d->where[at];
  const or_column* c = &p->t->cols[f->column_idx];
  switch (f->kind) {
    case TDBG_HAVING_CMP: {
      *first = at;
      tdbg_predicate leaf = {f->column_idx, f->type, f->op, c->validity != NULL, f->i64, f->f64};
      return pred_pass(&leaf, c, r);
    }
    case TDBG_HAVING_IS_NULL: *first = at; return !is_valid(c, r);
    case TDBG_HAVING_IS_NOT_NULL: *first = at; return is_valid(c, r);
    case TDBG_HAVING_NOT: return !where_eval(p, r, at - 1, first);
    default: {
      int is_and = f->kind == TDBG_HAVING_AND;
      int acc = is_and;
      int32_t end = at - 1;
      *first = at;
      for (uint32_t k = 0; k < f->arity; k++) {
        int32_t fi = end;
        int v = where_eval(p, r, end, &fi);
        acc = is_and ? (acc && v) : (acc || v);
        *first = fi;
        end = fi - 1;
      }
      return acc;
    }
  }
}

static int row_passes(const cplan* p, uint64_t r) {
  const tdbg_plan_desc* d = p->d;
  for (uint32_t k = 0; k < d->n_preds; k++)
    if (!pred_pass(&d->preds[k], &p->t->cols[d->preds[k].column_idx], r)) return 0;
  int32_t first = 0;
  return d->n_where == 0 || where_eval(p, r, (int32_t)d->n_where - 1, &first);
}

/* ---- one sub-batch: foldBatch (v2_lowcard_group.zig:977-1050) --------------- */
static void fold_batch(wstate* w, const cplan* p, uint64_t lo, uint64_t hi) {
  const or_table* t = p->t;
  uint64_t n = 0;
  for (uint64_t r = lo; r < hi; r++)
    if (row_passes(p, r)) w->sel[n++] = r;
  if (n == 0) return;
  /* packKeysForPart / packIntPart */
  for (uint64_t i = 0; i < n; i++) w->bkeys[i] = pack_key(p, w->sel[i]);
  /* runEnds */
  uint32_t runs = 0;
  for (uint64_t i = 1; i < n; i++)
    if (w->bkeys[i] != w->bkeys[i - 1]) w->ends[runs++] = (uint32_t)i;
  w->ends[runs++] = (uint32_t)n;
  /* gid per run: direct array or table (resolveRunsDirect / groupRows) */
  uint32_t s = 0;
  for (uint32_t ri = 0; ri < runs; ri++) {
    u128 key = w->bkeys[s];
    uint32_t gid;
    if (p->direct) {
      uint64_t at = ((uint64_t)key - p->direct_base) & p->direct_mask;
      int in_range = at < p->direct_span;
      if (in_range && w->direct_gids[at] != EMPTY_GID) {
        gid = w->direct_gids[at];
      } else {
        gid = ws_get_or_add(w, p, key);
        if (in_range) w->direct_gids[at] = gid;
      }
    } else {
      gid = ws_get_or_add(w, p, key);
    }
    if (w->err) return;
    w->run_gids[ri] = gid;
    w->counts[gid] += w->ends[ri] - s;
    s = w->ends[ri];
  }
  /* fold kernels, one aggregate at a time over the runs */
  for (uint32_t a = 0; a < p->n_aggs; a++) {
    const agg_plan* ag = &p->aggs[a];
    if (ag->op == OP_COUNT_STAR || ag->op == OP_COUNT_DISTINCT) continue; /* distinct: after the merge */
    const or_column* c = &t->cols[ag->col];
    uint32_t lo_i = 0;
    for (uint32_t ri = 0; ri < runs; ri++) {
      uint32_t hi_i = w->ends[ri];
      size_t o = (size_t)w->run_gids[ri] * p->n_aggs + a;
      if (ag->op == OP_COUNT_COL) {
        uint64_t cnt = 0;
        for (uint32_t i = lo_i; i < hi_i; i++) cnt += is_valid(c, w->sel[i]);
        w->ns[o] += cnt;
      } else if (ag->op == OP_SUM || ag->op == OP_AVG) {
        uint64_t cnt = 0;
        if (ag->is_float) {
          double acc = 0;
          for (uint32_t i = lo_i; i < hi_i; i++) {
            uint64_t r = w->sel[i];
            if (!is_valid(c, r)) continue;
            double v = load_f(c, r);
            acc += v;
            neumaier_add(&w->csum[o], &w->ccomp[o], v);
            cnt++;
          }
          if (cnt) { w->ns[o] += cnt; w->fslots[o] = w->fslots[o] + acc; }
        } else {
          i128 acc = 0;
          for (uint32_t i = lo_i; i < hi_i; i++) {
            uint64_t r = w->sel[i];
            if (!is_valid(c, r)) continue;
            acc += load_i(c, r);
            cnt++;
          }
          if (cnt) { w->ns[o] += cnt; w->slots[o] += acc; }
        }
      } else { /* MIN / MAX: first value seeds (was_empty) */
        int is_min = ag->op == OP_MIN;
        uint64_t cnt = 0;
        if (ag->is_float) {
          double best = 0;
          for (uint32_t i = lo_i; i < hi_i; i++) {
            uint64_t r = w->sel[i];
            if (!is_valid(c, r)) continue;
            double v = load_f(c, r);
            if (cnt == 0 || (is_min ? v < best : v > best)) best = v;
            cnt++;
          }
          if (cnt) {
            int was_empty = w->ns[o] == 0;
            w->ns[o] += cnt;
            int better = is_min ? best < w->fslots[o] : best > w->fslots[o];
            if (was_empty || better) w->fslots[o] = best;
          }
        } else {
          i128 best = 0;
          for (uint32_t i = lo_i; i < hi_i; i++) {
            uint64_t r = w->sel[i];
            if (!is_valid(c, r)) continue;
            i128 v = load_i(c, r);
            if (cnt == 0 || (is_min ? v < best : v > best)) best = v;
            cnt++;
          }
          if (cnt) {
            int was_empty = w->ns[o] == 0;
            w->ns[o] += cnt;
            int better = is_min ? best < w->slots[o] : best > w->slots[o];
            if (was_empty || better) w->slots[o] = best;
          }
        }
      }
      lo_i = hi_i;
    }
  }
}

/* ---- workers: tiles of TILE_RGS row groups from a shared cursor ------------- */
typedef struct {
  const cplan* p;
  wstate* w;
  _Atomic uint64_t* cursor;
  uint64_t n_tiles;
} worker_arg;

static void* worker_main(void* arg) {
  worker_arg* a = arg;
  const cplan* p = a->p;
  const or_table* t = p->t;
  uint64_t tile_rows = (uint64_t)t->rg_rows * TILE_RGS;
  for (;;) {
    uint64_t tile = atomic_fetch_add(a->cursor, 1);
    if (tile >= a->n_tiles) break;
    uint64_t lo = tile * tile_rows, hi = lo + tile_rows;
    if (hi > t->n_rows) hi = t->n_rows;
    /* each row group is scanned in autoScanBatch sub-batches */
    for (uint64_t rg = lo; rg < hi; rg += t->rg_rows) {
      uint64_t rg_end = rg + t->rg_rows < hi ? rg + t->rg_rows : hi;
      for (uint64_t b = rg; b < rg_end; b += p->batch_rows) {
        uint64_t be = b + p->batch_rows < rg_end ? b + p->batch_rows : rg_end;
        fold_batch(a->w, p, b, be);
        if (a->w->err) return NULL;
      }
    }
  }
  return NULL;
}

/* ---- result ---------------------------------------------------------------- */
struct or_result {
  uint64_t n_rows, n_groups;
  uint32_t n_cols;
  tdbg_column_view cols[MAX_KEYS + MAX_AGGS];
  double* comp[MAX_KEYS + MAX_AGGS];
};

uint64_t oracle_result_num_rows(const or_result* r) { return r->n_rows; }
uint64_t oracle_result_num_groups(const or_result* r) { return r->n_groups; }
uint32_t oracle_result_num_columns(const or_result* r) { return r->n_cols; }
int32_t oracle_result_column(const or_result* r, uint32_t i, tdbg_column_view* out) {
  if (i >= r->n_cols) return TDBG_ERR_INVALID;
  *out = r->cols[i];
  return 0;
}
int32_t oracle_result_compensated(const or_result* r, uint32_t i, double* out) {
  if (i >= r->n_cols || !r->comp[i]) return TDBG_ERR_INVALID;
  memcpy(out, r->comp[i], r->n_rows * sizeof(double));
  return 0;
}
void oracle_result_free(or_result* r) {
  if (!r) return;
  for (uint32_t i = 0; i < r->n_cols; i++) {
    free((void*)r->cols[i].data);
    free((void*)r->cols[i].validity);
    free(r->comp[i]);
  }
  free(r);
}

/* ---- sort (SortCtx.lessThan / compareBySpec) --------------------------------- */
typedef struct {
  const cplan* p;
  const wstate* m;
} sort_ctx;
static __thread const sort_ctx* g_sort;

static int float_order(double a, double b) { /* types.floatOrder: NaN largest */
  int an = isnan(a), bn = isnan(b);
  if (an || bn) return an && bn ? 0 : (an ? 1 : -1);
  return a < b ? -1 : (a > b ? 1 : 0);
}

static int part_is_null(u128 key, const key_part* kp) { return kp->nullable && !((key >> kp->valid_bit) & 1); }

static int64_t unpack_part(u128 key, const key_part* kp) {
  uint64_t raw = (uint64_t)(key >> kp->offset);
  if (kp->width >= 64) return (int64_t)raw;
  uint32_t sh = 64 - kp->width;
  return ((int64_t)(raw << sh)) >> sh;
}

/* One ORDER BY item, ascending: compareOrderKey / compareAggregateValue
 * (src/exec/v2_shape_group_topn.zig:1337-1395).  A NULL key part, and an
 * all-NULL SUM/AVG/MIN/MAX (aggIsNull :1411-1423: a nullable input with no
 * non-null value), orders first; the caller flips the result for DESC, so
 * it is last there.  Values compare as emitted: a BIGINT SUM by its wrapped
 * value (integerAggValue :1446-1452), a DECIMAL SUM as i128, AVG as
 * f64(sum)/f64(n)/10^scale (avgValue :1436-1444; a FLOAT/DOUBLE input sum/n),
 * floats under types.floatOrder (src/types.zig:579-587: -0 = +0, NaN last,
 * as the rank prefix OrderPrefix.ofFloat :1219-1226 orders it). */
/* Planted mistakes (tests only: each must make a comparator test fail, so
 * the tests that pin this comparator are shown to catch it).  0 = none. */
enum {
  PLANT_NONE = 0,
  PLANT_DECIMAL_LOW_WORD = 1, /* a DECIMAL128 SUM compared by its low 64 bits */
  PLANT_AVG_TRUNC = 2,        /* an integer AVG compared by sum / n in integer division */
  PLANT_NULL_AS_ZERO = 3,     /* an all-NULL aggregate compared as a zero value */
  PLANT_TIES_BY_GID = 4,      /* ties broken by first appearance instead of the packed key */
  PLANT_NAN_FIRST = 5,        /* NaN ordered below every number */
  PLANT_TIES_SIGNED = 6       /* ties broken by the key parts' signed values, not their bits */
};
static int g_plant = PLANT_NONE;
void oracle_plant(int which) { g_plant = which; }

static int compare_spec(const tdbg_order_spec* sp, uint32_t a, uint32_t b) {
  const cplan* p = g_sort->p;
  const wstate* m = g_sort->m;
  if (sp->target < p->n_parts) {
    const key_part* kp = &p->parts[sp->target];
    int an = part_is_null(m->keys[a], kp), bn = part_is_null(m->keys[b], kp);
    if (an != bn) return an ? -1 : 1;
    if (an) return 0;
    int64_t x = unpack_part(m->keys[a], kp);
    int64_t y = unpack_part(m->keys[b], kp);
    return x < y ? -1 : (x > y ? 1 : 0);
  }
  uint32_t i = sp->target - p->n_parts;
  const agg_plan* ag = &p->aggs[i];
  size_t oa = (size_t)a * p->n_aggs + i, ob = (size_t)b * p->n_aggs + i;
  if ((ag->op == OP_SUM || ag->op == OP_AVG || ag->op == OP_MIN || ag->op == OP_MAX) && g_plant != PLANT_NULL_AS_ZERO) {
    int an = ag->nullable && m->ns[oa] == 0, bn = ag->nullable && m->ns[ob] == 0;
    if (an || bn) return an && bn ? 0 : (an ? -1 : 1);
  }
  switch (ag->op) {
    case OP_COUNT_STAR: return m->counts[a] < m->counts[b] ? -1 : (m->counts[a] > m->counts[b]);
    case OP_COUNT_COL: case OP_COUNT_DISTINCT: return m->ns[oa] < m->ns[ob] ? -1 : (m->ns[oa] > m->ns[ob]);
    case OP_SUM: case OP_MIN: case OP_MAX:
      if (ag->is_float) {
        if (g_plant == PLANT_NAN_FIRST && (isnan(m->fslots[oa]) || isnan(m->fslots[ob])))
          return isnan(m->fslots[oa]) && isnan(m->fslots[ob]) ? 0 : (isnan(m->fslots[oa]) ? -1 : 1);
        return float_order(m->fslots[oa], m->fslots[ob]);
      }
      if (g_plant == PLANT_DECIMAL_LOW_WORD && ag->out_type == TDBG_DECIMAL128) {
        uint64_t x = (uint64_t)m->slots[oa], y = (uint64_t)m->slots[ob];
        return x < y ? -1 : (x > y);
      }
      if (ag->out_type == TDBG_BIGINT) {
        int64_t x = (int64_t)(uint64_t)m->slots[oa], y = (int64_t)(uint64_t)m->slots[ob];
        return x < y ? -1 : (x > y);
      }
      return m->slots[oa] < m->slots[ob] ? -1 : (m->slots[oa] > m->slots[ob]);
    default: { /* AVG, as emitted */
      uint64_t an = m->ns[oa], bn = m->ns[ob];
      if (g_plant == PLANT_AVG_TRUNC && !ag->is_float && an && bn) {
        i128 x = m->slots[oa] / (i128)an, y = m->slots[ob] / (i128)bn;
        return x < y ? -1 : (x > y);
      }
      double sc = pow(10.0, (double)ag->scale);
      double af = an == 0 ? 0 : (ag->is_float ? m->fslots[oa] / (double)an : (double)m->slots[oa] / (double)an / sc);
      double bf = bn == 0 ? 0 : (ag->is_float ? m->fslots[ob] / (double)bn : (double)m->slots[ob] / (double)bn / sc);
      return float_order(af, bf);
    }
  }
}

/* The silo's packed group key compared as a u128 (the last word of
 * rowOrderBefore, v2_shape_group_topn.zig:1327-1335).  Its layout
 * (validateShape :2051-2094, genericKeyFromViews silo_group_core.zig:
 * 5265-5283): part i at the bit offset of the widths before it, the value's
 * two's-complement bits masked to the type's width (intTypeBits :2698-2706),
 * a nullable part with a validity bit above them (0 = NULL, payload 0).  So
 * the comparison runs over the parts from the last (most significant) to
 * the first, each unsigned over (valid, bits).  Keys wider than 128 bits
 * are hashed by the silo, whose tie order is then unspecified. */
static int silo_key_cmp(const cplan* p, u128 ka, u128 kb) {
  for (int k = (int)p->n_parts - 1; k >= 0; k--) {
    const key_part* kp = &p->parts[k];
    int an = part_is_null(ka, kp), bn = part_is_null(kb, kp);
    if (an != bn) return an ? -1 : 1;
    if (an) continue;
    if (g_plant == PLANT_TIES_SIGNED) {
      int64_t x = unpack_part(ka, kp), y = unpack_part(kb, kp);
      if (x != y) return x < y ? -1 : 1;
      continue;
    }
    uint64_t mask = kp->width >= 64 ? ~0ull : ((1ull << kp->width) - 1);
    uint64_t x = (uint64_t)unpack_part(ka, kp) & mask, y = (uint64_t)unpack_part(kb, kp) & mask;
    if (x != y) return x < y ? -1 : 1;
  }
  return 0;
}

/* rowOrderBefore: the ORDER BY items, then the packed group key ascending,
 * so no two groups tie. */
static int gid_cmp(const void* pa, const void* pb) {
  uint32_t a = *(const uint32_t*)pa, b = *(const uint32_t*)pb;
  const tdbg_plan_desc* d = g_sort->p->d;
  for (uint32_t s = 0; s < d->n_order; s++) {
    int o = compare_spec(&d->order[s], a, b);
    if (o < 0) return d->order[s].desc ? 1 : -1;
    if (o > 0) return d->order[s].desc ? -1 : 1;
  }
  int k = g_plant == PLANT_TIES_BY_GID ? 0 : silo_key_cmp(g_sort->p, g_sort->m->keys[a], g_sort->m->keys[b]);
  if (k) return k;
  return a < b ? -1 : (a > b);
}

/* ---- HAVING (havingPasses / havingOutputNum) --------------------------------- */
typedef struct {
  i128 i;
  double f;
  int is_f;
  int null;
} hnum;

/* An output column's value for a group, as emitted (havingOutputNum). */
static hnum output_num(const cplan* p, const wstate* m, uint32_t target, uint32_t gid) {
  hnum v = {0, 0.0, 0, 0};
  if (target < p->n_parts) {
    if (part_is_null(m->keys[gid], &p->parts[target])) v.null = 1; /* keyPartIsNull */
    else v.i = unpack_part(m->keys[gid], &p->parts[target]);
    return v;
  }
  uint32_t a = target - p->n_parts;
  const agg_plan* ag = &p->aggs[a];
  size_t o = (size_t)gid * p->n_aggs + a;
  uint64_t ns = m->ns[o];
  if (ag->op == OP_COUNT_STAR) { v.i = (i128)m->counts[gid]; return v; }
  if (ag->op == OP_COUNT_COL || ag->op == OP_COUNT_DISTINCT) { v.i = (i128)ns; return v; }
  if (ag->nullable && ns == 0) { v.null = 1; return v; } /* aggIsNull */
  if (ag->op == OP_AVG) {
    v.is_f = 1;
    if (ag->is_float) v.f = ns ? m->fslots[o] / (double)ns : 0.0;
    else v.f = ns ? (double)m->slots[o] / (double)ns / pow(10.0, (double)ag->scale) : 0.0;
    return v;
  }
  if (ag->is_float) { v.is_f = 1; v.f = m->fslots[o]; return v; }
  v.i = ag->out_type == TDBG_BIGINT ? (i128)(int64_t)(uint64_t)m->slots[o] : m->slots[o]; /* integerAggValue */
  return v;
}

static int num_order(hnum a, hnum b) { /* numOrder */
  if (a.is_f || b.is_f) return float_order(a.is_f ? a.f : (double)a.i, b.is_f ? b.f : (double)b.i);
  return a.i < b.i ? -1 : (a.i > b.i);
}

/* Evaluates the subtree whose root is node `at` of the postfix array; its
 * first node's index goes to *first. */
static int having_eval(const cplan* p, const wstate* m, uint32_t gid, int32_t at, int32_t* first) {
  const tdbg_having_node* h = &p->d->having[at];
  switch (h->kind) {
    case TDBG_HAVING_CMP: {
      *first = at;
      hnum lhs = output_num(p, m, h->target, gid);
      if (lhs.null) return 0;
      hnum rhs = {0, 0.0, 0, 0};
      if (h->lit_float) { rhs.is_f = 1; rhs.f = h->lit_f64; }
      else rhs.i = (i128)(((u128)(uint64_t)h->lit_hi << 64) | (uint64_t)h->lit_lo);
      int o = num_order(lhs, rhs);
      switch (h->op) {
        case TDBG_EQ: return o == 0;
        case TDBG_NEQ: return o != 0;
        case TDBG_LT: return o < 0;
        case TDBG_LTE: return o <= 0;
        case TDBG_GT: return o > 0;
        default: return o >= 0;
      }
    }
    case TDBG_HAVING_IS_NULL: *first = at; return output_num(p, m, h->target, gid).null;
    case TDBG_HAVING_IS_NOT_NULL: *first = at; return !output_num(p, m, h->target, gid).null;
    case TDBG_HAVING_NOT: return !having_eval(p, m, gid, at - 1, first);
    default: { /* AND / OR: children end at at-1, each preceding the next */
      int is_and = h->kind == TDBG_HAVING_AND;
      int acc = is_and;
      int32_t end = at - 1;
      *first = at;
      for (uint32_t c = 0; c < h->arity; c++) {
        int32_t f = end;
        int r = having_eval(p, m, gid, end, &f);
        acc = is_and ? (acc && r) : (acc || r);
        *first = f;
        end = f - 1;
      }
      return acc;
    }
  }
}

/* ---- emit (emitGroup / appendInt / appendFloat) ------------------------------ */
static void put_int(void* data, uint32_t type, uint64_t row, i128 v) {
  switch (type) {
    case TDBG_BOOLEAN: ((uint8_t*)data)[row] = (uint8_t)v; break;
    case TDBG_TINYINT: ((int8_t*)data)[row] = (int8_t)v; break;
    case TDBG_SMALLINT: ((int16_t*)data)[row] = (int16_t)v; break;
    case TDBG_INT: case TDBG_DATE: ((int32_t*)data)[row] = (int32_t)v; break;
    case TDBG_DECIMAL128: ((i128*)data)[row] = v; break;
    case TDBG_DOUBLE: ((double*)data)[row] = (double)v; break;
    default: ((int64_t*)data)[row] = (int64_t)(uint64_t)v; break; /* @truncate */
  }
}

/* The best k gids of v[0..n) under gid_cmp, kept in place at v[0..kept)
 * (a heap whose top is the worst kept). */
typedef struct {
  uint32_t* v;
  uint64_t n, k, kept;
  const sort_ctx* ctx;  /* the comparator's context (thread-local: set in each worker) */
} topk_arg;

static void* topk_main(void* arg) {
  topk_arg* a = arg;
  g_sort = a->ctx;
  uint32_t* h = a->v;
  uint64_t n = 0;
  for (uint64_t i = 0; i < a->n; i++) {
    uint32_t g = a->v[i];
    if (n < a->k) {
      uint64_t c = n++;
      h[c] = g;
      while (c > 0 && gid_cmp(&h[(c - 1) / 2], &h[c]) < 0) {
        uint32_t x = h[c]; h[c] = h[(c - 1) / 2]; h[(c - 1) / 2] = x;
        c = (c - 1) / 2;
      }
    } else if (gid_cmp(&g, &h[0]) < 0) {
      h[0] = g;
      uint64_t c = 0;
      for (;;) {
        uint64_t l = 2 * c + 1, r = l + 1, big = c;
        if (l < n && gid_cmp(&h[l], &h[big]) > 0) big = l;
        if (r < n && gid_cmp(&h[r], &h[big]) > 0) big = r;
        if (big == c) break;
        uint32_t x = h[c]; h[c] = h[big]; h[big] = x;
        c = big;
      }
    }
  }
  a->kept = n;
  return NULL;
}

/* The groups `live` (n_live gids; null: every gid below m->n_groups) of m,
 * HAVING-filtered, ordered, sliced and written as result columns. */
static int emit_live(const cplan* p, const wstate* m, const uint32_t* live, uint64_t n_live, or_result** out) {
  const tdbg_plan_desc* d = p->d;
  uint64_t ng = 0, ng_sorted = ~0ull;
  if (!live) n_live = m->n_groups;
  uint32_t* order = malloc((n_live + 1) * sizeof(uint32_t));
  if (!order) return -1;
  for (uint64_t j = 0; j < n_live; j++) {
    const uint32_t i = live ? live[j] : (uint32_t)j;
    int32_t first = 0;
    if (d->n_having == 0 || having_eval(p, m, i, (int32_t)d->n_having - 1, &first)) order[ng++] = i;
  }
  if (d->n_order > 0) {
    sort_ctx ctx = {p, m};
    g_sort = &ctx;
    const uint64_t k = d->limit ? d->offset + d->limit : 0;
    if (k && k < ng / 4) {
      /* only the first offset + limit groups of the order matter (the
       * reference's TopSet/RankedSet, v2_shape_group_topn.zig:846-870): each
       * of the workers keeps the best k of its share in a heap whose top is
       * the worst kept, and the kept ones are sorted; the comparator is a
       * total order, so they hold the full sort's first k */
      uint32_t nt = p->dop > 1 && ng > (1u << 20) ? p->dop : 1;
      if (nt > 64) nt = 64;
      topk_arg ta[64];
      pthread_t th[64];
      for (uint32_t i = 0; i < nt; i++) {
        ta[i] = (topk_arg){order + ng * i / nt, ng * (i + 1) / nt - ng * i / nt, k, 0, &ctx};
        if (nt > 1) pthread_create(&th[i], NULL, topk_main, &ta[i]);
        else topk_main(&ta[i]);
      }
      uint64_t n = 0;
      for (uint32_t i = 0; i < nt; i++) {
        if (nt > 1) pthread_join(th[i], NULL);
        memmove(order + n, ta[i].v, ta[i].kept * sizeof(uint32_t));  /* shares are in order: no overlap */
        n += ta[i].kept;
      }
      qsort(order, n, sizeof(uint32_t), gid_cmp);
      if (n > k) n = k;
      ng_sorted = n;
    } else {
      qsort(order, ng, sizeof(uint32_t), gid_cmp);
    }
    g_sort = NULL;
  }
  uint64_t start = d->offset < ng ? d->offset : ng;
  uint64_t end = d->limit == 0 ? ng : (start + d->limit < ng ? start + d->limit : ng);
  if (end > ng_sorted) end = ng_sorted;  /* (a top-k selection kept exactly these) */
  uint64_t nr = end - start;
  or_result* r = calloc(1, sizeof *r);
  if (!r) { free(order); return -1; }
  r->n_rows = nr;
  r->n_groups = ng;
  r->n_cols = p->n_parts + p->n_aggs;
  for (uint32_t k = 0; k < p->n_parts; k++) {
    tdbg_column_view* c = &r->cols[k];
    c->type = p->parts[k].type;
    c->n_rows = nr;
    c->nullable = p->parts[k].nullable;
    void* data = calloc(nr + 1, type_bytes(c->type));
    uint8_t* valid = calloc(nr / 8 + 1, 1);
    int any_null = 0;
    for (uint64_t i = 0; i < nr; i++) {
      u128 key = m->keys[order[start + i]];
      if (part_is_null(key, &p->parts[k])) { any_null = 1; continue; }
      put_int(data, c->type, i, unpack_part(key, &p->parts[k]));
      valid[i >> 3] |= (uint8_t)(1u << (i & 7));
    }
    c->data = data;
    if (any_null) c->validity = valid; else free(valid);
  }
  for (uint32_t a = 0; a < p->n_aggs; a++) {
    const agg_plan* ag = &p->aggs[a];
    tdbg_column_view* c = &r->cols[p->n_parts + a];
    c->type = ag->out_type;
    c->scale = (ag->out_type == TDBG_DECIMAL128 || ag->out_type == TDBG_DECIMAL64) ? ag->scale : 0;
    c->nullable = !(ag->op == OP_COUNT_STAR || ag->op == OP_COUNT_COL || ag->op == OP_COUNT_DISTINCT);
    c->n_rows = nr;
    void* data = calloc(nr + 1, type_bytes(c->type));
    uint8_t* valid = calloc(nr / 8 + 1, 1);
    double* comp = NULL;
    int want_comp = ag->is_float && (ag->op == OP_SUM || ag->op == OP_AVG);
    if (want_comp) comp = calloc(nr + 1, sizeof(double));
    int any_null = 0;
    for (uint64_t i = 0; i < nr; i++) {
      uint32_t gid = order[start + i];
      size_t o = (size_t)gid * p->n_aggs + a;
      uint64_t ns = m->ns[o];
      int is_null = 0;
      switch (ag->op) {
        case OP_COUNT_STAR: ((int64_t*)data)[i] = (int64_t)m->counts[gid]; break;
        case OP_COUNT_COL: case OP_COUNT_DISTINCT: ((int64_t*)data)[i] = (int64_t)ns; break;
        case OP_SUM: case OP_MIN: case OP_MAX:
          if (ns == 0) { is_null = 1; break; }
          if (ag->is_float) {
            if (ag->out_type == TDBG_FLOAT) ((float*)data)[i] = (float)m->fslots[o];
            else ((double*)data)[i] = m->fslots[o];
            if (comp) comp[i] = m->csum[o] + m->ccomp[o];
          } else {
            put_int(data, ag->out_type, i, m->slots[o]);
          }
          break;
        default: /* AVG */
          if (ns == 0) { is_null = 1; break; }
          if (ag->is_float) {
            ((double*)data)[i] = m->fslots[o] / (double)ns;
            if (comp) comp[i] = (m->csum[o] + m->ccomp[o]) / (double)ns;
          } else {
            ((double*)data)[i] = (double)m->slots[o] / (double)ns / pow(10.0, (double)ag->scale);
          }
          break;
      }
      if (is_null) any_null = 1; else valid[i >> 3] |= (uint8_t)(1u << (i & 7));
    }
    c->data = data;
    if (any_null) c->validity = valid; else free(valid);
    r->comp[p->n_parts + a] = comp;
  }
  free(order);
  *out = r;
  return 0;
}

static int emit(const cplan* p, const wstate* m, or_result** out) { return emit_live(p, m, NULL, 0, out); }

/* ---- COUNT(DISTINCT): the (group, value) composites of the filtered rows
 * with a non-NULL value, deduplicated, counted per group (the distinct sets
 * of foldDistinctValues/compositeOf, v2_lowcard_group.zig:1359-1418,
 * :1497-1499, counted after the merge :1538-1552) ------------------------------ */
typedef struct {
  uint32_t gid;
  int64_t v;
} dpair;

static int dpair_cmp(const void* pa, const void* pb) {
  const dpair* a = pa;
  const dpair* b = pb;
  if (a->gid != b->gid) return a->gid < b->gid ? -1 : 1;
  return a->v < b->v ? -1 : (a->v > b->v);
}

static uint32_t ws_find(const wstate* w, const cplan* p, u128 key) {
  uint64_t mask = w->cap - 1;
  for (uint64_t h = hash_key(p, key) & mask; w->tgids[h] != EMPTY_GID; h = (h + 1) & mask)
    if (w->tkeys[h] == key) return w->tgids[h];
  return EMPTY_GID;
}

static int distinct_counts(const cplan* p, wstate* m) {
  const or_table* t = p->t;
  for (uint32_t a = 0; a < p->n_aggs; a++) {
    const agg_plan* ag = &p->aggs[a];
    if (ag->op != OP_COUNT_DISTINCT) continue;
    const or_column* c = &t->cols[ag->col];
    dpair* pairs = malloc((t->n_rows + 1) * sizeof(dpair));
    if (!pairs) return -1;
    uint64_t n = 0;
    for (uint64_t r = 0; r < t->n_rows; r++) {
      if (!is_valid(c, r) || !row_passes(p, r)) continue;
      uint32_t gid = ws_find(m, p, pack_key(p, r));
      if (gid == EMPTY_GID) { free(pairs); return -1; } /* every passing row has its group */
      pairs[n].gid = gid;
      if (ag->is_float) { /* canonicalFloatBits of the f64 (types.zig:593-600, silo_group_core.zig:8052): one +0, one NaN */
        double x = load_f(c, r);
        uint64_t bits;
        if (isnan(x)) bits = 0x7ff8000000000000ull;
        else if (x == 0.0) bits = 0;
        else memcpy(&bits, &x, 8);
        pairs[n].v = (int64_t)bits;
      } else {
        pairs[n].v = load_i(c, r);
      }
      n++;
    }
    qsort(pairs, n, sizeof(dpair), dpair_cmp);
    for (uint64_t i = 0; i < n; i++)
      if (i == 0 || pairs[i].gid != pairs[i - 1].gid || pairs[i].v != pairs[i - 1].v)
        m->ns[(size_t)pairs[i].gid * p->n_aggs + a]++;
    free(pairs);
  }
  return 0;
}

/* ---- merge (mergeScalars) ---------------------------------------------------- */
static int merge_into(wstate* m, const cplan* p, const wstate* w) {
  for (uint64_t g = 0; g < w->n_groups; g++) {
    uint32_t mg = ws_get_or_add(m, p, w->keys[g]);
    if (m->err) return -1;
    m->counts[mg] += w->counts[g];
    for (uint32_t i = 0; i < p->n_aggs; i++) {
      const agg_plan* ag = &p->aggs[i];
      size_t wo = (size_t)g * p->n_aggs + i, mo = (size_t)mg * p->n_aggs + i;
      uint64_t w_ns = w->ns[wo];
      if (ag->op == OP_COUNT_STAR) w_ns = w->counts[g];
      if (w_ns == 0) continue;
      uint64_t had = m->ns[mo];
      m->ns[mo] += w_ns;
      switch (ag->op) {
        case OP_SUM: case OP_AVG:
          if (ag->is_float) {
            m->fslots[mo] = m->fslots[mo] + w->fslots[wo];
            neumaier_add(&m->csum[mo], &m->ccomp[mo], w->csum[wo]);
            m->ccomp[mo] += w->ccomp[wo];
          } else {
            m->slots[mo] += w->slots[wo];
          }
          break;
        case OP_MIN:
          if (ag->is_float) { if (had == 0 || w->fslots[wo] < m->fslots[mo]) m->fslots[mo] = w->fslots[wo]; }
          else if (had == 0 || w->slots[wo] < m->slots[mo]) m->slots[mo] = w->slots[wo];
          break;
        case OP_MAX:
          if (ag->is_float) { if (had == 0 || w->fslots[wo] > m->fslots[mo]) m->fslots[mo] = w->fslots[wo]; }
          else if (had == 0 || w->slots[wo] > m->slots[mo]) m->slots[mo] = w->slots[wo];
          break;
        default: break;
      }
    }
  }
  return 0;
}

/* ---- plan compile ------------------------------------------------------------ */
static int compile(cplan* p, const tdbg_plan_desc* d, const or_table* t) {
  memset(p, 0, sizeof *p);
  p->t = t;
  p->d = d;
  if (d->n_keys == 0 || d->n_keys > MAX_KEYS) { set_err("key count"); return TDBG_ERR_UNSUPPORTED; }
  if (d->n_aggs == 0 || d->n_aggs > MAX_AGGS) { set_err("agg count"); return TDBG_ERR_UNSUPPORTED; }
  uint32_t bits = 0;
  for (uint32_t k = 0; k < d->n_keys; k++) {
    const tdbg_key_part* kp = &d->keys[k];
    if (kp->column_idx >= t->n_cols) { set_err("key column"); return TDBG_ERR_INVALID; }
    if (t->cols[kp->column_idx].type != kp->type) { set_err("key type"); return TDBG_ERR_INVALID; }
    int w = key_type_bits(kp->type);
    if (!w) { set_err("key type not integer"); return TDBG_ERR_UNSUPPORTED; }
    p->parts[k] = (key_part){kp->column_idx, kp->type, bits, (uint32_t)w, t->cols[kp->column_idx].validity != NULL, 0};
    bits += (uint32_t)w;
  }
  /* a nullable part's validity bit sits above every value bit (packValidity,
   * v2_lowcard_group.zig:954-966; tryBuild :245-272) */
  for (uint32_t k = 0; k < d->n_keys; k++)
    if (p->parts[k].nullable) p->parts[k].valid_bit = bits++;
  if (bits > 128) { set_err("key wider than 128 bits"); return TDBG_ERR_UNSUPPORTED; }
  p->n_parts = d->n_keys;
  p->key_bits = bits;
  uint64_t row_bytes = 0;
  for (uint32_t k = 0; k < d->n_keys; k++) row_bytes += type_bytes(d->keys[k].type);
  for (uint32_t a = 0; a < d->n_aggs; a++) {
    const tdbg_agg_spec* s = &d->aggs[a];
    agg_plan* ag = &p->aggs[a];
    ag->col = s->column_idx;
    ag->in_type = s->input_type;
    ag->scale = s->input_scale;
    if (s->op == TDBG_AGG_COUNT_STAR) { ag->op = OP_COUNT_STAR; ag->out_type = TDBG_BIGINT; ag->col = -1; continue; }
    if (s->column_idx < 0 || (uint32_t)s->column_idx >= t->n_cols) { set_err("agg column"); return TDBG_ERR_INVALID; }
    if (t->cols[s->column_idx].type != s->input_type) { set_err("agg type"); return TDBG_ERR_INVALID; }
    ag->nullable = t->cols[s->column_idx].validity != NULL;
    row_bytes += type_bytes(s->input_type);
    if (s->op == TDBG_AGG_COUNT) { ag->op = OP_COUNT_COL; ag->out_type = TDBG_BIGINT; continue; }
    if (s->op == TDBG_AGG_COUNT_DISTINCT) { /* distinctValueBits: an int's bits, a float's canonical f64 bits */
      if (!key_type_bits(s->input_type) && !is_float_type(s->input_type)) {
        set_err("COUNT(DISTINCT) input type");
        return TDBG_ERR_UNSUPPORTED;
      }
      ag->op = OP_COUNT_DISTINCT;
      ag->is_float = is_float_type(s->input_type);
      ag->out_type = TDBG_BIGINT;
      continue;
    }
    int fl = is_float_type(s->input_type);
    if (!fl && !key_type_bits(s->input_type)) { set_err("agg input type"); return TDBG_ERR_UNSUPPORTED; }
    if ((s->input_type == TDBG_DATE || s->input_type == TDBG_DATETIME) &&
        (s->op == TDBG_AGG_SUM || s->op == TDBG_AGG_AVG)) { set_err("temporal SUM/AVG"); return TDBG_ERR_UNSUPPORTED; }
    ag->is_float = fl;
    switch (s->op) {
      case TDBG_AGG_SUM:
        ag->op = OP_SUM;
        ag->out_type = fl ? TDBG_DOUBLE : (s->input_type == TDBG_DECIMAL64 ? TDBG_DECIMAL128 : TDBG_BIGINT);
        break;
      case TDBG_AGG_AVG: ag->op = OP_AVG; ag->out_type = TDBG_DOUBLE; break;
      case TDBG_AGG_MIN: ag->op = OP_MIN; ag->out_type = s->input_type; break;
      case TDBG_AGG_MAX: ag->op = OP_MAX; ag->out_type = s->input_type; break;
      default: set_err("agg op"); return TDBG_ERR_UNSUPPORTED;
    }
  }
  p->n_aggs = d->n_aggs;
  for (uint32_t i = 0; i < d->n_preds; i++) {
    if (d->preds[i].column_idx >= t->n_cols) { set_err("predicate column"); return TDBG_ERR_INVALID; }
    if (d->preds[i].op > TDBG_GTE) { set_err("predicate op"); return TDBG_ERR_INVALID; }
  }
  for (uint32_t i = 0; i < d->n_order; i++)
    if (d->order[i].target >= d->n_keys + d->n_aggs) { set_err("order target"); return TDBG_ERR_INVALID; }
  /* WHERE tree: a postfix array that forms exactly one tree */
  if (d->n_where && !d->where) { set_err("where array"); return TDBG_ERR_INVALID; }
  int64_t wdepth = 0;
  for (uint32_t i = 0; i < d->n_where; i++) {
    const tdbg_filter_node* f = &d->where[i];
    if (f->kind <= TDBG_HAVING_IS_NOT_NULL) {
      if (f->column_idx >= t->n_cols || (f->kind == TDBG_HAVING_CMP && f->op > TDBG_GTE)) {
        set_err("where leaf");
        return TDBG_ERR_INVALID;
      }
      wdepth++;
    } else if (f->kind == TDBG_HAVING_NOT) {
      if (wdepth < 1) { set_err("where NOT"); return TDBG_ERR_INVALID; }
    } else if (f->kind == TDBG_HAVING_AND || f->kind == TDBG_HAVING_OR) {
      if ((int64_t)f->arity > wdepth) { set_err("where arity"); return TDBG_ERR_INVALID; }
      wdepth -= (int64_t)f->arity - 1;
    } else {
      set_err("where kind");
      return TDBG_ERR_INVALID;
    }
  }
  if (d->n_where && wdepth != 1) { set_err("where shape"); return TDBG_ERR_INVALID; }
  /* HAVING: a postfix array that forms exactly one tree */
  if (d->n_having && !d->having) { set_err("having array"); return TDBG_ERR_INVALID; }
  if (d->n_having > TDBG_MAX_HAVING) { set_err("having size"); return TDBG_ERR_UNSUPPORTED; }
  int64_t depth = 0;
  for (uint32_t i = 0; i < d->n_having; i++) {
    const tdbg_having_node* h = &d->having[i];
    if (h->kind <= TDBG_HAVING_IS_NOT_NULL) {
      if (h->target >= d->n_keys + d->n_aggs || (h->kind == TDBG_HAVING_CMP && h->op > TDBG_GTE)) {
        set_err("having leaf");
        return TDBG_ERR_INVALID;
      }
      depth++;
    } else if (h->kind == TDBG_HAVING_NOT) {
      if (depth < 1) { set_err("having NOT"); return TDBG_ERR_INVALID; }
    } else if (h->kind == TDBG_HAVING_AND || h->kind == TDBG_HAVING_OR) {
      if ((int64_t)h->arity > depth) { set_err("having arity"); return TDBG_ERR_INVALID; }
      depth -= (int64_t)h->arity - 1;
    } else {
      set_err("having kind");
      return TDBG_ERR_INVALID;
    }
  }
  if (d->n_having && depth != 1) { set_err("having shape"); return TDBG_ERR_INVALID; }
  /* autoScanBatch (exec.zig:77-85) */
  uint64_t raw = row_bytes ? ((1u << 20) / 8) / row_bytes : 32768;
  if (raw < 2048) raw = 2048;
  if (raw > 32768) raw = 32768;
  p->batch_rows = (uint32_t)(raw & ~63ull);
  /* directRangeOf (:509-530) over the key column's min/max */
  if (d->n_keys == 1 && t->n_rows > 0 && !p->parts[0].nullable) { /* a nullable part has no direct range (:510) */
    const or_column* c = &t->cols[d->keys[0].column_idx];
    int64_t mn = load_i(c, 0), mx = mn;
    for (uint64_t r = 1; r < t->n_rows; r++) {
      int64_t v = load_i(c, r);
      if (v < mn) mn = v;
      if (v > mx) mx = v;
    }
    uint64_t mask = p->parts[0].width >= 64 ? ~0ull : ((1ull << p->parts[0].width) - 1);
    uint64_t base = (uint64_t)mn & mask;
    uint64_t span = (((uint64_t)mx - (uint64_t)mn) & mask) + 1;
    uint64_t est = d->est_groups ? d->est_groups : span;
    if (span != 0 && span <= DIRECT_MAX_SPAN && span <= DIRECT_SPAN_PER_GROUP * (est ? est : 1)) {
      p->direct = 1;
      p->direct_base = base;
      p->direct_mask = mask;
      p->direct_span = (uint32_t)span;
    }
  }
  return 0;
}

int32_t oracle_group_by(const tdbg_plan_desc* d, const or_table* t, uint32_t dop, or_result** out) {
  cplan p;
  int rc = compile(&p, d, t);
  if (rc) return rc;
  if (dop == 0) dop = 1;
  p.dop = dop;
  if (t->rg_rows == 0) { set_err("rg_rows"); return TDBG_ERR_INVALID; }
  uint64_t n_rgs = (t->n_rows + t->rg_rows - 1) / t->rg_rows;
  uint64_t n_tiles = (n_rgs + TILE_RGS - 1) / TILE_RGS;
  if (dop > n_tiles && n_tiles > 0) dop = (uint32_t)n_tiles;
  uint64_t est = d->est_groups ? d->est_groups : 1024;
  wstate* ws = calloc(dop, sizeof(wstate));
  worker_arg* args = calloc(dop, sizeof(worker_arg));
  pthread_t* th = calloc(dop, sizeof(pthread_t));
  _Atomic uint64_t cursor = 0;
  int fail = !ws || !args || !th;
  for (uint32_t i = 0; i < dop && !fail; i++) {
    if (ws_init(&ws[i], &p, est)) fail = 1;
    args[i] = (worker_arg){&p, &ws[i], &cursor, n_tiles};
  }
  if (!fail) {
    if (dop == 1) {
      worker_main(&args[0]);
    } else {
      for (uint32_t i = 0; i < dop; i++) pthread_create(&th[i], NULL, worker_main, &args[i]);
      for (uint32_t i = 0; i < dop; i++) pthread_join(th[i], NULL);
    }
    for (uint32_t i = 0; i < dop; i++) if (ws[i].err) fail = 1;
  }
  wstate merged;
  int have_merged = 0;
  if (!fail) {
    if (dop == 1) {
      merged = ws[0];
      memset(&ws[0], 0, sizeof ws[0]);
    } else {
      if (ws_init(&merged, &p, est)) fail = 1;
      for (uint32_t i = 0; i < dop && !fail; i++) if (merge_into(&merged, &p, &ws[i])) fail = 1;
    }
    have_merged = 1;
  }
  if (!fail) {
    /* COUNT(*) reads `counts`; the worker slots only track ns for inputs */
    if (distinct_counts(&p, &merged) || emit(&p, &merged, out)) fail = 1;
  }
  if (have_merged) ws_free(&merged);
  for (uint32_t i = 0; ws && i < dop; i++) ws_free(&ws[i]);
  free(ws); free(args); free(th);
  if (fail) { set_err("out of memory"); return TDBG_ERR_OOM; }
  return 0;
}

/* ---- the silo's high-cardinality route / bucket fold, restated for the CPU
 * baseline (silo_group_core.zig): rows are routed by routeHashKey of their
 * packed key (:5218-5224) into 256 buckets (bucketIndexHash, power of two:
 * the hash's low bits, :5232-5235), staged per worker with their aggregate
 * inputs (the stage/scatter of :9228-9400), then each bucket is re-split by
 * the next hash bits into parts of a few thousand rows (the deferred radix
 * parts, :10799-10830) and each part folded into a table of its own
 * (foldDeferredBucket, :10868-10960).  Buckets hold disjoint keys, so their
 * groups are concatenated, not merged.  The result is the same as
 * oracle_group_by's (same folds, same emit); only the work is arranged
 * differently, for timing.  COUNT(DISTINCT) declines here. -------------------- */

#define RX_BUCKETS 256u
#define RX_PART_ROWS 4096u

typedef struct {
  u128 key;
  uint64_t valid; /* bit a: aggregate a's input is non-NULL */
  uint64_t v[];   /* per aggregate: the input (i64, or f64 bits) */
} rx_row;

typedef struct {
  uint8_t* data;
  uint64_t n, cap;
} rx_vec;

static inline uint64_t route_hash(u128 key) {
  uint64_t lo = (uint64_t)key, hi = (uint64_t)(key >> 64);
  uint64_t x = lo ^ (hi * 0x9e3779b97f4a7c15ull);
  x ^= x >> 32;
  return x;
}

typedef struct {
  const cplan* p;
  _Atomic uint64_t* cursor;
  uint64_t n_tiles, stride;
  rx_vec* b; /* [RX_BUCKETS] this worker's staged rows per bucket */
  uint64_t* sel;
  int err;
} rx_stage_arg;

static void* rx_stage_main(void* arg) {
  rx_stage_arg* a = arg;
  const cplan* p = a->p;
  const or_table* t = p->t;
  const uint64_t tile_rows = (uint64_t)t->rg_rows * TILE_RGS;
  for (;;) {
    uint64_t tile = atomic_fetch_add(a->cursor, 1);
    if (tile >= a->n_tiles) break;
    uint64_t lo = tile * tile_rows, hi = lo + tile_rows < t->n_rows ? lo + tile_rows : t->n_rows;
    for (uint64_t r = lo; r < hi; r++) {
      if (!row_passes(p, r)) continue;
      const u128 key = pack_key(p, r);
      rx_vec* v = &a->b[route_hash(key) & (RX_BUCKETS - 1)];
      if (v->n == v->cap) {
        uint64_t nc = v->cap ? v->cap * 2 : 1024;
        uint8_t* nd = realloc(v->data, nc * a->stride);
        if (!nd) { a->err = 1; return NULL; }
        v->data = nd;
        v->cap = nc;
      }
      rx_row* row = (rx_row*)(v->data + v->n++ * a->stride);
      row->key = key;
      row->valid = 0;
      for (uint32_t g = 0; g < p->n_aggs; g++) {
        const agg_plan* ag = &p->aggs[g];
        row->v[g] = 0;
        if (ag->col < 0) continue;
        const or_column* c = &t->cols[ag->col];
        if (!is_valid(c, r)) continue;
        row->valid |= 1ull << g;
        if (ag->is_float) { double x = load_f(c, r); memcpy(&row->v[g], &x, 8); }
        else row->v[g] = (uint64_t)load_i(c, r);
      }
    }
  }
  return NULL;
}

/* One row folded into group `gid` of `w` (the fold kernels of fold_batch,
 * one row at a time: MIN/MAX seeded by the group's first value). */
static void rx_fold_row(wstate* w, const cplan* p, uint32_t gid, const rx_row* row) {
  w->counts[gid]++;
  for (uint32_t a = 0; a < p->n_aggs; a++) {
    const agg_plan* ag = &p->aggs[a];
    if (ag->op == OP_COUNT_STAR || !((row->valid >> a) & 1)) continue;
    size_t o = (size_t)gid * p->n_aggs + a;
    const int first = w->ns[o] == 0;
    w->ns[o]++;
    if (ag->op == OP_COUNT_COL) continue;
    if (ag->is_float) {
      double x;
      memcpy(&x, &row->v[a], 8);
      if (ag->op == OP_SUM || ag->op == OP_AVG) {
        w->fslots[o] += x;
        neumaier_add(&w->csum[o], &w->ccomp[o], x);
      } else if (first || (ag->op == OP_MIN ? x < w->fslots[o] : x > w->fslots[o])) {
        w->fslots[o] = x;
      }
    } else {
      const i128 x = (int64_t)row->v[a];
      if (ag->op == OP_SUM || ag->op == OP_AVG) w->slots[o] += x;
      else if (first || (ag->op == OP_MIN ? x < w->slots[o] : x > w->slots[o])) w->slots[o] = x;
    }
  }
}

typedef struct {
  const cplan* p;
  rx_stage_arg* st; /* the stage workers (their bucket vectors) */
  uint32_t n_st;
  uint64_t stride;
  _Atomic uint32_t* cursor;
  wstate* all;             /* the query's group arrays, one slot per staged row */
  const uint64_t* base;    /* [RX_BUCKETS] bucket b's groups start at gid base[b] */
  uint64_t* n_groups;      /* [RX_BUCKETS] groups bucket b found */
  int err;
} rx_fold_arg;

static void* rx_fold_main(void* arg) {
  rx_fold_arg* a = arg;
  const cplan* p = a->p;
  const uint64_t stride = a->stride;
  uint8_t* tmp = NULL;
  uint64_t tmp_cap = 0;
  for (;;) {
    uint32_t b = atomic_fetch_add(a->cursor, 1);
    if (b >= RX_BUCKETS) break;
    uint64_t n = 0;
    for (uint32_t w = 0; w < a->n_st; w++) n += a->st[w].b[b].n;
    /* the bucket's groups go to gids base[b].. of the shared arrays (a bucket
     * has at most as many groups as rows): a view of them */
    wstate gv;
    memset(&gv, 0, sizeof gv);
    wstate* g = &gv;
    const uint64_t o0 = a->base[b], na = p->n_aggs;
    g->keys = a->all->keys + o0;
    g->counts = a->all->counts + o0;
    g->slots = a->all->slots + o0 * na;
    g->fslots = a->all->fslots + o0 * na;
    g->csum = a->all->csum + o0 * na;
    g->ccomp = a->all->ccomp + o0 * na;
    g->ns = a->all->ns + o0 * na;
    g->gcap = n;
    a->n_groups[b] = 0;
    if (n == 0) continue;
    /* the deferred parts: the bucket's rows split by the next hash bits */
    uint32_t parts = 1;
    while ((uint64_t)parts * RX_PART_ROWS < n && parts < 4096) parts *= 2;
    if (tmp_cap < n) {
      free(tmp);
      tmp = malloc(n * stride);
      tmp_cap = tmp ? n : 0;
      if (!tmp) { a->err = 1; break; }
    }
    uint64_t* off = calloc(parts + 1, sizeof(uint64_t));
    if (!off) { a->err = 1; break; }
    for (uint32_t w = 0; w < a->n_st; w++) {
      const rx_vec* v = &a->st[w].b[b];
      for (uint64_t i = 0; i < v->n; i++)
        off[((route_hash(((const rx_row*)(v->data + i * stride))->key) >> 8) & (parts - 1)) + 1]++;
    }
    for (uint32_t q = 0; q < parts; q++) off[q + 1] += off[q];
    for (uint32_t w = 0; w < a->n_st; w++) {  /* worker order, row order within a worker */
      const rx_vec* v = &a->st[w].b[b];
      for (uint64_t i = 0; i < v->n; i++) {
        const rx_row* row = (const rx_row*)(v->data + i * stride);
        const uint32_t q = (route_hash(row->key) >> 8) & (parts - 1);
        memcpy(tmp + off[q]++ * stride, row, stride);
      }
    }
    /* off[q] now ends part q; fold each part into a fresh table */
    uint64_t s = 0;
    for (uint32_t q = 0; q < parts && !a->err; q++) {
      const uint64_t e = off[q];
      uint64_t cap = 16;
      while (cap < 2 * (e - s)) cap *= 2;
      u128* tk = malloc(cap * sizeof(u128));
      uint32_t* tg = malloc(cap * sizeof(uint32_t));
      if (!tk || !tg) { free(tk); free(tg); a->err = 1; break; }
      for (uint64_t i = 0; i < cap; i++) tg[i] = EMPTY_GID;
      for (uint64_t i = s; i < e; i++) {
        const rx_row* row = (const rx_row*)(tmp + i * stride);
        uint64_t h = hash_key(p, row->key) & (cap - 1);
        while (tg[h] != EMPTY_GID && tk[h] != row->key) h = (h + 1) & (cap - 1);
        if (tg[h] == EMPTY_GID) {
          const uint32_t gid = (uint32_t)g->n_groups++;  /* < n: every new group has a row */
          g->keys[gid] = row->key;
          g->counts[gid] = 0;
          for (uint32_t x = 0; x < p->n_aggs; x++) {
            size_t o = (size_t)gid * p->n_aggs + x;
            g->slots[o] = 0; g->fslots[o] = 0; g->csum[o] = 0; g->ccomp[o] = 0; g->ns[o] = 0;
          }
          tk[h] = row->key;
          tg[h] = gid;
        }
        rx_fold_row(g, p, tg[h], row);
      }
      free(tk);
      free(tg);
      s = e;
    }
    free(off);
    a->n_groups[b] = g->n_groups;
  }
  free(tmp);
  return NULL;
}

int32_t oracle_group_by_radix(const tdbg_plan_desc* d, const or_table* t, uint32_t dop, or_result** out) {
  cplan p;
  int rc = compile(&p, d, t);
  if (rc) return rc;
  for (uint32_t a = 0; a < p.n_aggs; a++)
    if (p.aggs[a].op == OP_COUNT_DISTINCT) { set_err("COUNT(DISTINCT) on the radix baseline"); return TDBG_ERR_UNSUPPORTED; }
  if (p.n_aggs > 60) { set_err("agg count"); return TDBG_ERR_UNSUPPORTED; }
  if (dop == 0) dop = 1;
  p.dop = dop;
  if (t->rg_rows == 0) { set_err("rg_rows"); return TDBG_ERR_INVALID; }
  const uint64_t stride = sizeof(rx_row) + 8ull * p.n_aggs;
  const uint64_t n_rgs = (t->n_rows + t->rg_rows - 1) / t->rg_rows;
  const uint64_t n_tiles = (n_rgs + TILE_RGS - 1) / TILE_RGS;
  int fail = 0;
  /* stage: workers route their tiles' rows into per-worker buckets */
  rx_stage_arg* st = calloc(dop, sizeof *st);
  pthread_t* th = calloc(dop, sizeof *th);
  _Atomic uint64_t tcur = 0;
  if (!st || !th) fail = 1;
  for (uint32_t w = 0; w < dop && !fail; w++) {
    st[w] = (rx_stage_arg){&p, &tcur, n_tiles, stride, calloc(RX_BUCKETS, sizeof(rx_vec)), NULL, 0};
    if (!st[w].b) fail = 1;
  }
  if (!fail) {
    for (uint32_t w = 0; w < dop; w++) pthread_create(&th[w], NULL, rx_stage_main, &st[w]);
    for (uint32_t w = 0; w < dop; w++) pthread_join(th[w], NULL);
    for (uint32_t w = 0; w < dop; w++) if (st[w].err) fail = 1;
  }
  /* fold: workers take whole buckets, writing their groups into one set of
   * arrays with a slot per staged row (bucket b from base[b]) */
  uint64_t base[RX_BUCKETS + 1], ngr[RX_BUCKETS];
  base[0] = 0;
  for (uint32_t b = 0; b < RX_BUCKETS; b++) {
    uint64_t n = 0;
    for (uint32_t w = 0; !fail && w < dop; w++) n += st[w].b[b].n;
    base[b + 1] = base[b] + n;
  }
  wstate m;
  memset(&m, 0, sizeof m);
  const uint64_t slots = base[RX_BUCKETS] + 1, na = p.n_aggs ? p.n_aggs : 1;
  if (!fail) {
    m.keys = malloc(slots * sizeof(u128));
    m.counts = malloc(slots * sizeof(uint64_t));
    m.slots = malloc(slots * na * sizeof(i128));
    m.fslots = malloc(slots * na * sizeof(double));
    m.csum = malloc(slots * na * sizeof(double));
    m.ccomp = malloc(slots * na * sizeof(double));
    m.ns = malloc(slots * na * sizeof(uint64_t));
    if (!m.keys || !m.counts || !m.slots || !m.fslots || !m.csum || !m.ccomp || !m.ns) fail = 1;
  }
  rx_fold_arg* fa = fail ? NULL : calloc(dop, sizeof *fa);
  _Atomic uint32_t bcur = 0;
  if (!fail && !fa) fail = 1;
  if (!fail) {
    for (uint32_t w = 0; w < dop; w++) fa[w] = (rx_fold_arg){&p, st, dop, stride, &bcur, &m, base, ngr, 0};
    for (uint32_t w = 0; w < dop; w++) pthread_create(&th[w], NULL, rx_fold_main, &fa[w]);
    for (uint32_t w = 0; w < dop; w++) pthread_join(th[w], NULL);
    for (uint32_t w = 0; w < dop; w++) if (fa[w].err) fail = 1;
  }
  /* the buckets' groups (disjoint keys) as one list of gids, then emitted */
  if (!fail) {
    uint64_t total = 0;
    for (uint32_t b = 0; b < RX_BUCKETS; b++) total += ngr[b];
    uint32_t* live = malloc((total + 1) * sizeof(uint32_t));
    if (!live || base[RX_BUCKETS] >= EMPTY_GID) fail = 1;
    uint64_t at = 0;
    for (uint32_t b = 0; b < RX_BUCKETS && !fail; b++)
      for (uint64_t i = 0; i < ngr[b]; i++) live[at++] = (uint32_t)(base[b] + i);
    m.n_groups = slots;
    if (!fail && emit_live(&p, &m, live, total, out)) fail = 1;
    free(live);
  }
  free(m.keys); free(m.counts); free(m.slots); free(m.fslots); free(m.csum); free(m.ccomp); free(m.ns);
  free(fa);
  for (uint32_t w = 0; st && w < dop; w++) {
    for (uint32_t b = 0; st[w].b && b < RX_BUCKETS; b++) free(st[w].b[b].data);
    free(st[w].b);
  }
  free(st);
  free(th);
  if (fail) { set_err("out of memory"); return TDBG_ERR_OOM; }
  return 0;
}
