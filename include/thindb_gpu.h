/*
 * thindb_gpu.h — C-ABI of the MI355X GROUP BY engine (libthindb_gpu.so).
 *
 * The library replaces thinDB's CPU grouped operators for the shape
 *   scan -> filter -> GROUP BY (COUNT/SUM/AVG/MIN/MAX) -> ORDER BY ... LIMIT
 * behind the engine's `exec.Query` operator contract.  The caller (thinDB's
 * planner, or any FFI host) keeps decoded column blocks resident in HBM with
 * `tdbg_column_stage`, describes one grouped query with `tdbg_plan_desc`
 * and runs it with `tdbg_execute`, which returns ONE result batch: group
 * keys first, then aggregates, ordered and sliced by ORDER BY / OFFSET /
 * LIMIT exactly as the reference operators emit it.
 *
 * Reference interfaces each entry point stands in for (paths relative to
 * the thinDB source tree):
 *   - handshake: `abi_version()` + i32 status, as the dlopen handshake in
 *     src/api/zigfn.zig:1-26, :165-183, :598-603;
 *   - column blocks: the decoded-block cache keyed by
 *     (table_uid, segment_id, row_group_idx, column_idx),
 *     src/storage/cache.zig:59-70, filled by
 *     src/storage/segment_reader.zig:115 (decodeColumnMaybeCached);
 *   - plan: the grouped request src/exec/v2_shape_group_topn.zig:63-101
 *     (`Request`) with src/exec/aggregate.zig:237-268 (`AggSpec`),
 *     src/exec/predicate.zig:33-51 (`Predicate`, `PredicateOp`) and
 *     src/exec/sort.zig:28-31 (`SortSpec`);
 *   - execute: the first `next()` of src/exec/v2_lowcard_group.zig:1775-1784
 *     and src/exec/v2_shape_group_topn.zig:620-646, which run the whole
 *     query and return one `Batch` (src/exec/exec.zig:175-247);
 *   - result columns: `ColumnView` (src/storage/column.zig:27-60), validity
 *     bit 1 = valid, bit r at byte r>>3, bit r&7.
 *
 * Status codes (all entry points returning int32_t): 0 = ok, negative =
 * error; `tdbg_last_error()` gives the calling thread's last message.
 *   TDBG_ERR_UNSUPPORTED  -> the caller declines (returns null from its
 *                            tryBuild) and runs the CPU operator instead;
 *   TDBG_ERR_OOM          -> error.MemoryBudgetExceeded (src/memory.zig:31-35);
 *   TDBG_ERR_CANCELLED    -> error.QueryCancelled (src/exec/exec.zig:913-921);
 *   anything else         -> error.Unexpected with tdbg_last_error().
 * No entry point ever returns a partial result.
 *
 * Thread safety: every entry point may be called from many threads at once
 * (thinDB runs one connection thread per query).  One context owns one GPU.
 */
#ifndef THINDB_GPU_H
#define THINDB_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDBG_ABI_VERSION 3u

enum tdbg_status {
  TDBG_OK = 0,
  TDBG_PENDING = 1,            /* tdbg_poll: not finished yet */
  TDBG_ERR_INVALID = -1,       /* malformed argument or plan */
  TDBG_ERR_UNSUPPORTED = -2,   /* shape not covered: decline to the CPU path */
  TDBG_ERR_OOM = -3,           /* device memory / admission budget exhausted */
  TDBG_ERR_CANCELLED = -4,     /* tdbg_cancel() observed */
  TDBG_ERR_DEVICE = -5,        /* HIP runtime or device failure */
  TDBG_ERR_NOT_STAGED = -6     /* a referenced column block is not resident */
};

/* Physical column types; mirrors the integer/float subset of
 * src/types.zig (DECIMAL(p<=18,s) is an i64 mantissa, types.zig:377-383;
 * DATETIME is i64 microseconds and DATE i32 days, DESIGN.md:58-75). */
enum tdbg_type {
  TDBG_BOOLEAN = 1,    /* u8 0/1 */
  TDBG_TINYINT = 2,    /* i8 */
  TDBG_SMALLINT = 3,   /* i16 */
  TDBG_INT = 4,        /* i32 */
  TDBG_BIGINT = 5,     /* i64 */
  TDBG_DATE = 6,       /* i32 days */
  TDBG_DATETIME = 7,   /* i64 microseconds */
  TDBG_DECIMAL64 = 8,  /* i64 mantissa, scale in the spec */
  TDBG_FLOAT = 9,      /* f32 */
  TDBG_DOUBLE = 10,    /* f64 */
  TDBG_DECIMAL128 = 11 /* i128 mantissa; result columns only (SUM of DECIMAL) */
};

/* src/exec/aggregate.zig:88-92 (count, sum, min, max, avg); COUNT(*) is
 * count with no column (aggregate.zig:244-246).  COUNT(DISTINCT col)
 * (aggregate.zig:130 count_distinct; planned by v2_lowcard_group.zig:363-387,
 * folded by :1348-1418) counts each group's distinct non-NULL values of an
 * integer-family or float column (a float by its canonical f64 bits: one
 * +0.0, one NaN; types.zig:593-600): BIGINT, never NULL (0 when every value
 * is NULL). */
enum tdbg_agg_op {
  TDBG_AGG_COUNT_STAR = 0,
  TDBG_AGG_COUNT = 1,
  TDBG_AGG_SUM = 2,
  TDBG_AGG_AVG = 3,
  TDBG_AGG_MIN = 4,
  TDBG_AGG_MAX = 5,
  TDBG_AGG_COUNT_DISTINCT = 6
};

/* src/exec/predicate.zig:33 `PredicateOp = enum { eq, neq, lt, lte, gt, gte }`. */
enum tdbg_pred_op {
  TDBG_EQ = 0,
  TDBG_NEQ = 1,
  TDBG_LT = 2,
  TDBG_LTE = 3,
  TDBG_GT = 4,
  TDBG_GTE = 5
};

typedef struct tdbg_context tdbg_context;
typedef struct tdbg_plan tdbg_plan;
typedef struct tdbg_result tdbg_result;
typedef struct tdbg_batch tdbg_batch;
typedef struct tdbg_ticket tdbg_ticket;

/* Identity of one decoded column block: src/storage/cache.zig:59-70 (`Key`). */
typedef struct tdbg_column_key {
  uint64_t table_uid;
  uint64_t segment_id;
  uint32_t row_group_idx;
  uint32_t column_idx;
} tdbg_column_key;

/* Zone-map statistics of one block (src/storage/format.zig:265-270).  When
 * `has_minmax` is 0 the library computes them from the host buffer. */
typedef struct tdbg_block_stats {
  int64_t min_i, max_i;  /* integer-family columns */
  double min_f, max_f;   /* FLOAT / DOUBLE columns */
  uint32_t has_minmax;
  uint32_t pad;
} tdbg_block_stats;

/* One surviving row group of the scan (src/exec/rg_tiles.zig:1-30). */
typedef struct tdbg_rg_ref {
  uint64_t segment_id;
  uint32_t row_group_idx;
  uint32_t pad;
} tdbg_rg_ref;

/* One GROUP BY key part (packKeysForPart, src/exec/v2_lowcard_group.zig:886-966;
 * src/exec/v2_shape_group_topn.zig:2051-2094).  A key column staged with a
 * validity bitmap is a nullable part: its NULL rows form their own group
 * (apart from the value 0), emitted with a NULL key, ordered first ascending
 * and last descending (compareBySpec, v2_lowcard_group.zig:2219-2224). */
typedef struct tdbg_key_part {
  uint32_t column_idx;
  uint32_t type; /* enum tdbg_type, integer family only */
} tdbg_key_part;

/* One aggregate: src/exec/aggregate.zig:237-268 (`AggSpec`), resolved to
 * the input column's physical type. */
typedef struct tdbg_agg_spec {
  uint32_t op;          /* enum tdbg_agg_op */
  int32_t column_idx;   /* -1 for COUNT(*) */
  uint32_t input_type;  /* enum tdbg_type of the input column */
  uint8_t input_scale;  /* DECIMAL scale, else 0 */
  uint8_t nullable;     /* input column carries a validity bitmap */
  uint8_t pad[2];
} tdbg_agg_spec;

/* One `col op literal` conjunct of the WHERE (predicate.zig:35-51).  The
 * literal is already coerced to the column's type: `i64` for the integer
 * family, `f64` for FLOAT/DOUBLE.  A NULL value fails every comparison. */
typedef struct tdbg_predicate {
  uint32_t column_idx;
  uint32_t type;     /* enum tdbg_type of the column */
  uint32_t op;       /* enum tdbg_pred_op */
  uint32_t nullable; /* column carries a validity bitmap */
  int64_t i64;
  double f64;
} tdbg_predicate;

/* ORDER BY item (src/exec/sort.zig:28-31).  `target` indexes the output
 * columns: [0, n_keys) are the keys, [n_keys, n_keys + n_aggs) the aggs. */
typedef struct tdbg_order_spec {
  uint32_t target;
  uint32_t desc;
} tdbg_order_spec;

/* HAVING (`Request.having_filter`, src/exec/v2_shape_group_topn.zig:77;
 * evaluator havingPasses/havingOutputNum :2461-2530): the predicate tree over
 * the output columns in postfix order, applied to every group after
 * aggregation and before ORDER BY / OFFSET / LIMIT.
 *
 *   CMP          pushes `output[target] op literal`; values and literals
 *                are numbers as valueAsNum (:2391-2407) makes them: an
 *                integer (`lit_lo`/`lit_hi`, i128 little-endian halves) or,
 *                with `lit_float`, a double; a comparison with a double on
 *                either side is made in f64 under types.floatOrder
 *                (numOrder :2409-2422).  An output is the emitted value
 *                (a BIGINT SUM wrapped, AVG divided by 10^scale, a FLOAT
 *                MIN/MAX before narrowing to f32); a NULL output makes the
 *                leaf false.
 *   IS_NULL / IS_NOT_NULL  push whether `output[target]` is NULL (or not).
 *   NOT          replaces the top with its negation (two-valued: NOT of a
 *                leaf over a NULL is true, as in the reference).
 *   AND / OR     pop `arity` entries and push their conjunction /
 *                disjunction (arity 0: true / false).
 * The array must leave exactly one entry; at most TDBG_MAX_HAVING nodes. */
#define TDBG_MAX_HAVING 32u
enum tdbg_having_kind {
  TDBG_HAVING_CMP = 0,
  TDBG_HAVING_IS_NULL = 1,
  TDBG_HAVING_IS_NOT_NULL = 2,
  TDBG_HAVING_NOT = 3,
  TDBG_HAVING_AND = 4,
  TDBG_HAVING_OR = 5,
};
typedef struct tdbg_having_node {
  uint32_t kind;      /* enum tdbg_having_kind */
  uint32_t target;    /* CMP / IS_[NOT_]NULL: output column index */
  uint32_t op;        /* CMP: enum tdbg_pred_op */
  uint32_t arity;     /* AND / OR: children popped */
  uint32_t lit_float; /* CMP: 1 = the literal is `lit_f64` */
  uint32_t pad;
  int64_t lit_lo;     /* CMP integer literal, low 64 bits */
  int64_t lit_hi;     /* its high 64 bits (sign extension for an i64) */
  double lit_f64;
} tdbg_having_node;

/* A WHERE beyond a conjunction of `col op literal` leaves
 * (`PredicateExpr` .leaf / .is_null / .is_not_null / .not / .and / .or,
 * src/exec/predicate.zig:55-160), in postfix order, evaluated per row as
 * evaluatePredicate does (:1727-1790): a comparison with a NULL value does
 * not match, IS [NOT] NULL reads the validity, NOT inverts its child's
 * matches, AND / OR combine `arity` children (arity 0: every row / none).
 * Kinds are enum tdbg_having_kind; a comparison's literal is coerced to the
 * column's type as in tdbg_predicate.  The plan's filter is the AND of
 * `preds` and this tree.  At most TDBG_MAX_WHERE nodes. */
#define TDBG_MAX_WHERE 12u
typedef struct tdbg_filter_node {
  uint32_t kind;       /* enum tdbg_having_kind */
  uint32_t column_idx; /* CMP / IS_[NOT_]NULL */
  uint32_t type;       /* CMP: enum tdbg_type of the column */
  uint32_t op;         /* CMP: enum tdbg_pred_op */
  uint32_t arity;      /* AND / OR: children popped */
  uint32_t pad;
  int64_t i64;         /* CMP literal, integer family */
  double f64;          /* CMP literal, FLOAT / DOUBLE */
} tdbg_filter_node;

/* A grouped query: the part of v2_shape_group_topn.zig:63-101 (`Request`)
 * the GPU operator consumes.  `limit == 0` means no limit, as in
 * src/exec/v2_lowcard_group.zig:2072-2081.  The WHERE is the AND of `preds`
 * and the `where` tree; `having` (n_having nodes, may be NULL when 0) filters
 * the groups. */
typedef struct tdbg_plan_desc {
  uint64_t table_uid;
  uint32_t n_keys;
  uint32_t n_aggs;
  const tdbg_key_part* keys;
  const tdbg_agg_spec* aggs;
  uint32_t n_preds;
  uint32_t n_order;
  const tdbg_predicate* preds;
  const tdbg_order_spec* order;
  uint64_t limit;
  uint64_t offset;
  uint64_t est_groups; /* planner's NDV-product bound; 0 = unknown */
  uint32_t n_having;
  uint32_t n_where;
  const tdbg_having_node* having;
  const tdbg_filter_node* where; /* n_where nodes (may be NULL when 0) */
} tdbg_plan_desc;

/* One result column, library-owned host memory valid until
 * tdbg_result_free().  DECIMAL128 values are 16-byte little-endian two's
 * complement.  `validity` is NULL when no row is NULL. */
typedef struct tdbg_column_view {
  uint32_t type;  /* enum tdbg_type */
  uint32_t scale; /* DECIMAL scale, else 0 */
  uint32_t nullable;
  uint32_t pad;
  const void* data;
  const uint8_t* validity;
  uint64_t n_rows;
} tdbg_column_view;

/* Device-side timing of the last execution of a plan (HIP events on the
 * plan's stream), for measurement. */
typedef struct tdbg_exec_stats {
  double total_ms;      /* whole execution on the device */
  double scan_ms;       /* fused scan/filter/aggregate kernel(s) */
  uint64_t rows_scanned;
  uint64_t bytes_scanned; /* native-width bytes of every referenced column */
  uint64_t n_groups;
  uint32_t lds_mode;    /* 0 none, 1 direct, 2 hash */
  uint32_t global_mode; /* 1 dense, 2 hash, 3 radix-partitioned by key range, 4 hash-partitioned (wide/sparse keys) */
  uint32_t specialized; /* 1: plan-specialised kernel, 0: generic kernel */
  uint32_t lds_parts;   /* workgroups sharing one tile range (direct mode) */
} tdbg_exec_stats;

/* ---- handshake / context --------------------------------------------- */
uint32_t tdbg_abi_version(void);
const char* tdbg_last_error(void);
/* `hbm_budget_bytes` caps what the context allocates (0 = 90% of the
 * device); beyond it calls fail with TDBG_ERR_OOM. */
int32_t tdbg_context_create(int32_t device, uint64_t hbm_budget_bytes, tdbg_context** out);
void tdbg_context_destroy(tdbg_context* ctx);
uint64_t tdbg_context_bytes_in_use(const tdbg_context* ctx);

/* ---- HBM column cache (src/storage/cache.zig:1-8, :59-70) -------------- */
/* Copies one decoded block (`n_rows` values of `type`, plus an optional
 * validity bitmap) into HBM.  Blocks of one row group must agree on n_rows.
 * Segments are immutable, so a staged block never needs invalidation. */
int32_t tdbg_column_stage(tdbg_context* ctx, const tdbg_column_key* key, uint32_t type,
                          const void* host_values, const uint8_t* host_validity,
                          uint64_t n_rows, const tdbg_block_stats* stats);
/* Stages one block as the decompressed payload thinDB's block cache holds it
 * (src/storage/cache.zig:1-8): the validity bitmap first when `has_nulls`
 * ((n_rows + 7) / 8 bytes, bit 1 = valid), then the body of its encoding,
 * expanded in HBM by the library — decodeColumnPayload and its raw/FOR/RLE
 * cases (src/storage/segment_reader.zig:1727-1760, :1989-2075, :2133-2247):
 *   TDBG_ENC_RAW  n_rows native values
 *   TDBG_ENC_FOR  [base i128 LE][width u8 = 1|2|4|8][3 pad][n_rows codes of
 *                 `width` bytes LE]; value = base + code, wrapping at the
 *                 column's width (forBlockOf, :1809-1815)
 *   TDBG_ENC_RLE  [n_runs u32][value width u8 = the column's][3 pad]
 *                 [n_runs values][n_runs u32 lengths]; runs past n_rows are
 *                 clamped, rows past the last run read 0 (rleBlockOf,
 *                 :2098-2109)
 * Only the encoded body crosses PCIe.  FOR/RLE hold integer-family columns
 * (the writer encodes no others); dict/FSST blocks carry strings and are
 * TDBG_ERR_UNSUPPORTED.  `stats` as in tdbg_column_stage (the block's
 * footer Stats, format.zig:265-270); without them the library derives them
 * on the host. */
enum { TDBG_ENC_RAW = 0, TDBG_ENC_FOR = 1, TDBG_ENC_DICT = 2, TDBG_ENC_RLE = 3, TDBG_ENC_FSST = 4 };
int32_t tdbg_column_stage_encoded(tdbg_context* ctx, const tdbg_column_key* key, uint32_t type, uint32_t encoding,
                                  const void* payload, uint64_t payload_bytes, uint64_t n_rows, uint32_t has_nulls,
                                  const tdbg_block_stats* stats);
int32_t tdbg_column_evict(tdbg_context* ctx, const tdbg_column_key* key);

/* ---- segment files (src/storage/segment_reader.zig) --------------------- */
/* A thinDB segment file (.dat, format version 17) read from a host image of
 * the whole file (a read or an mmap; it must outlive the tdbg_segment).
 * tdbg_segment_open parses the 32-byte header, the trailer and the footer as
 * readSegment does (segment_reader.zig:1350-1455; string-table directory
 * :1459-1512) and fails TDBG_ERR_INVALID on anything malformed
 * (format.Error: BadMagic, CorruptFooter, ...) and TDBG_ERR_UNSUPPORTED on
 * another format version.  `n_columns` is the table schema's column count,
 * which the footer layout depends on. */
typedef struct tdbg_segment tdbg_segment;
typedef struct tdbg_segment_info {
  uint64_t segment_id;
  uint64_t row_count;
  uint64_t schema_fingerprint;
  uint32_t n_row_groups;
  uint32_t n_columns;
} tdbg_segment_info;
/* One column block's header (format.ColumnBlockHeader, format.zig:203-231)
 * and its footer Stats (format.zig:233-270: i128 min/max/sum as lo/hi
 * halves, null_count). */
typedef struct tdbg_segment_block {
  uint32_t row_count;
  uint32_t payload_bytes; /* uncompressed_size */
  uint32_t stored_bytes;  /* compressed_size */
  uint8_t compression;    /* 0 none, 1 zstd, 2 lz4 */
  uint8_t has_nulls;
  uint8_t int_streams;
  uint8_t encoding;       /* TDBG_ENC_*, or 5 = segment dictionary (strings) */
  uint64_t null_count;
  uint64_t min_lo, min_hi, max_lo, max_hi, sum_lo, sum_hi;
} tdbg_segment_block;
int32_t tdbg_segment_open(const void* bytes, uint64_t n_bytes, uint32_t n_columns, tdbg_segment** out);
void tdbg_segment_close(tdbg_segment* seg);
int32_t tdbg_segment_get_info(const tdbg_segment* seg, tdbg_segment_info* out);
int32_t tdbg_segment_block_info(const tdbg_segment* seg, uint32_t row_group_idx, uint32_t column_idx,
                                tdbg_segment_block* out);
/* The block's payload as thinDB's block cache holds it: decompressed (zstd or
 * LZ4 block, through the system's libzstd.so.1 / liblz4.so.1) and, for an
 * int_streams block, its integer arrays restored (readBlockRaw and
 * restoreIntStreams, segment_reader.zig:1620-1716; int_streams.zig).
 * `capacity` must hold payload_bytes.  Host only: no device is touched. */
int32_t tdbg_segment_read_block(const tdbg_segment* seg, uint32_t row_group_idx, uint32_t column_idx, void* out,
                                uint64_t capacity);
/* Reads one block as tdbg_segment_read_block and stages it with
 * tdbg_column_stage_encoded under {table_uid, the segment's id, row group,
 * column} (decodeColumnMaybeCached, segment_reader.zig:115, filling the HBM
 * column cache); an integer-family block with non-NULL rows carries its
 * footer zone map.  String blocks are TDBG_ERR_UNSUPPORTED. */
int32_t tdbg_segment_stage(tdbg_context* ctx, const tdbg_segment* seg, uint64_t table_uid, uint32_t row_group_idx,
                           uint32_t column_idx, uint32_t type);
/* Pre-sizes a table's HBM arenas for `rows` rows (a hint; arenas grow). */
int32_t tdbg_table_reserve(tdbg_context* ctx, uint64_t table_uid, uint64_t rows);
int32_t tdbg_table_drop(tdbg_context* ctx, uint64_t table_uid);

/* ---- plans and execution ----------------------------------------------- */
/* TDBG_ERR_UNSUPPORTED means "decline": the caller runs its CPU operator. */
int32_t tdbg_plan_create(tdbg_context* ctx, const tdbg_plan_desc* desc, tdbg_plan** out);
void tdbg_plan_destroy(tdbg_plan* plan);
/* Runs the query over the listed row groups (NULL = every staged row group
 * of the table, in staging order) and returns one result batch. */
int32_t tdbg_execute(tdbg_plan* plan, const tdbg_rg_ref* rgs, uint32_t n_rgs, tdbg_result** out);
int32_t tdbg_plan_last_stats(const tdbg_plan* plan, tdbg_exec_stats* out);
/* Asks every execution of `plan` already asked for - running, waiting for the
 * plan, or queued in a server - to stop; each returns TDBG_ERR_CANCELLED (a
 * served request cancelled while queued never reaches the GPU).  Executions
 * asked for after this call are not affected.  Cooperative, as
 * memory.checkCancelled (src/exec/exec.zig:913-921): device work already
 * launched finishes first. */
int32_t tdbg_cancel(tdbg_plan* plan);

/* ---- batched execution --------------------------------------------------
 * No counterpart in the reference, which runs each grouped query alone on
 * its connection thread (exec.Query per request, src/exec/exec.zig:1221-1239;
 * net/mysql/server.zig:191).  A batch runs many plans - each over every
 * staged row group of its table - as one persistent device launch (captured
 * once as a HIP graph and replayed), so small and uneven queries share the
 * GPU instead of paying a launch-and-drain each.  Plans with ORDER BY,
 * LIMIT, OFFSET or COUNT(DISTINCT) are declined (TDBG_ERR_UNSUPPORTED): run
 * them with tdbg_execute.  Every result is the one tdbg_execute returns for the plan.
 * A batch of 64 plans or more runs as 2 waves (THINDB_GPU_BATCH_WAVES sets
 * the number; at least 32 plans each) so that the result copy of each wave
 * overlaps the kernels of the next. */
typedef struct tdbg_batch_stats {
  double total_ms;      /* device time of the last execution, result copy included */
  double scan_ms;       /* the persistent scan/aggregate launch (0 when replayed as a graph) */
  uint64_t rows;        /* rows scanned over all plans */
  uint64_t bytes;       /* native-width bytes of every referenced column */
  uint64_t result_bytes;/* device-to-host bytes of the results */
  uint32_t n_plans;
  uint32_t n_items;     /* work items of the persistent launch */
  uint32_t graph;       /* 1: replayed as a captured HIP graph */
  uint32_t waves;       /* sub-batches run one after another, each one's result copy overlapping the
                           kernels of the next (0: one launch sequence) */
} tdbg_batch_stats;

/* How the batch runs one plan (chosen per (rows, groups) bucket by the
 * autotuner, else by the single-plan planner). */
typedef struct tdbg_batch_route {
  uint32_t strategy;    /* 1 LDS direct table, 2 HBM atomics, 3 LDS hash table, 4 radix partitions */
  uint32_t lds_parts;   /* direct: workgroups splitting the key range */
  uint32_t radix_bits;  /* radix: level-1 digit bits */
  uint32_t items;       /* work items of this plan */
} tdbg_batch_route;

int32_t tdbg_batch_create(tdbg_context* ctx, tdbg_plan* const* plans, uint32_t n_plans, tdbg_batch** out);
/* Measures the candidate strategies on the batch's own plans, bucket by
 * bucket of (rows, estimated groups), `reps` timed runs each, keeps the
 * fastest per bucket in the context (later batches reuse it) and re-plans. */
int32_t tdbg_batch_autotune(tdbg_batch* batch, uint32_t reps);
int32_t tdbg_batch_execute(tdbg_batch* batch);
/* Result of plan i of the last execution.  Its columns are views into host
 * memory the result keeps alive; they are not overwritten by later
 * executions of the batch. */
int32_t tdbg_batch_result(tdbg_batch* batch, uint32_t i, tdbg_result** out);
int32_t tdbg_batch_last_stats(const tdbg_batch* batch, tdbg_batch_stats* out);
int32_t tdbg_batch_plan_info(const tdbg_batch* batch, uint32_t i, tdbg_batch_route* out);
void tdbg_batch_destroy(tdbg_batch* batch);

/* ---- serving ----------------------------------------------------------------
 * No counterpart in the reference: thindb-server runs each query on its
 * connection thread (net/mysql/server.zig:191) with a process-wide CPU
 * scheduler (util/core_scheduler.zig:1-28).  A server is the device-side
 * request loop: connection threads submit plans (tdbg_server_submit replaces
 * a direct tdbg_execute) and wait on tickets; `lanes` worker threads each
 * keep one unit of work on the GPU (own stream), so units overlap.  A unit is
 * either one large request (many rows, or ORDER BY/LIMIT/OFFSET) run alone,
 * or the small requests queued at that moment (up to max_batch, FIFO) run as
 * one batched persistent launch: continuous batching.  With two lanes or
 * more, lane 0 runs small requests only, so large ones never hold them back.
 * Requests are admitted in arrival order while the estimated device bytes of
 * the running units (group tables, partition scratch, results) stay within
 * admit_bytes. */
typedef struct tdbg_server tdbg_server;
typedef struct tdbg_ticket tdbg_ticket;
typedef struct tdbg_server_config {
  uint32_t lanes;        /* units on the device at once (0 = 3) */
  uint32_t max_batch;    /* requests per batched launch (0 = 64) */
  uint64_t admit_bytes;  /* estimated device bytes of running units (0 = a quarter of the budget) */
  uint64_t large_rows;   /* requests scanning this many rows run alone (0 = 4M) */
} tdbg_server_config;
typedef struct tdbg_ticket_report {
  int64_t submit_ns;     /* steady-clock times of submission, dispatch and completion */
  int64_t start_ns;
  int64_t done_ns;
  uint64_t est_bytes;    /* admission estimate */
  uint32_t batch_size;   /* requests that shared its launch (1 = ran alone) */
  uint32_t lane;
  int32_t status;        /* TDBG_PENDING until done */
  uint32_t pad;
} tdbg_ticket_report;
typedef struct tdbg_server_counters {
  uint64_t submitted;
  uint64_t completed;
  uint64_t units;            /* launches (a batch counts once) */
  uint64_t batched;          /* requests that ran in batches of two or more */
  uint64_t peak_admitted;    /* highest estimated bytes of running units */
  uint64_t admission_waits;  /* times the queue head waited for memory */
} tdbg_server_counters;

int32_t tdbg_server_start(tdbg_context* ctx, const tdbg_server_config* cfg, tdbg_server** out);
/* Queues one execution of `plan` over every staged row group of its table.
 * The plan must outlive the ticket's completion. */
int32_t tdbg_server_submit(tdbg_server* server, tdbg_plan* plan, tdbg_ticket** out);
/* The same over the row groups `rgs` names (a query's snapshot, as the
 * operator passes them to tdbg_execute; n_rgs 0 scans nothing), copied. */
int32_t tdbg_server_submit_rgs(tdbg_server* server, tdbg_plan* plan, const tdbg_rg_ref* rgs, uint32_t n_rgs,
                               tdbg_ticket** out);
/* TDBG_PENDING while queued or running, else the execution's status. */
int32_t tdbg_ticket_poll(tdbg_ticket* ticket);
/* Blocks until done; on TDBG_OK hands over the result (caller frees it). */
int32_t tdbg_ticket_wait(tdbg_ticket* ticket, tdbg_result** out);
/* The same, giving up after about timeout_us microseconds with TDBG_PENDING
 * (the ticket stays valid): a connection thread waits in slices and checks
 * its KILL flag between them (memory.checkCancelled, src/memory.zig:60),
 * forwarding a KILL with tdbg_cancel. */
int32_t tdbg_ticket_wait_for(tdbg_ticket* ticket, uint64_t timeout_us, tdbg_result** out);
int32_t tdbg_ticket_info(const tdbg_ticket* ticket, tdbg_ticket_report* out);
void tdbg_ticket_free(tdbg_ticket* ticket);
int32_t tdbg_server_stats(const tdbg_server* server, tdbg_server_counters* out);
/* Runs what is queued, then stops the lanes. */
void tdbg_server_stop(tdbg_server* server);

/* ---- allocation-free execution ------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * or pinned memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Runs `plan` as tdbg_execute does and
 * returns the same rows, but calls no hipMalloc, hipFree, hipHostMalloc or
 * other allocating routine: every device buffer its kernels use comes from
 * one scratch array of TDBG_NOALLOC_SCRATCH_BYTES compiled into the library,
 * which each GPU holds until the process ends (its address is resolved by
 * tdbg_context_create).  The result is written into the caller's host buffer
 * `out` of `out_capacity` bytes (per output column: rows x the type's width,
 * then one byte per row for a nullable column, each 16-byte aligned; that
 * is at most rows x (width + 1) + 16 bytes per column); `cols` (n_cols = the
 * plan's output columns) receive views into `out` (validity as a bitmap,
 * NULL when no row is NULL) and *n_rows the row count.  Limits: an execution
 * whose buffers exceed the scratch array fails with TDBG_ERR_UNSUPPORTED (run
 * it with tdbg_execute); `out` too small fails with TDBG_ERR_INVALID.  The
 * array keeps the last call's contents; every kernel initialises what it
 * reads.  Calls on one GPU share its array, so they are serialised by the
 * library (a second call waits for the first).  The plan-specialised scan
 * kernel is used when an earlier tdbg_execute of the same plan shape
 * compiled it (compiling allocates), else the generic one. */
#define TDBG_NOALLOC_SCRATCH_BYTES (64ull << 20)
int32_t tdbg_execute_noalloc(tdbg_plan* plan, const tdbg_rg_ref* rgs, uint32_t n_rgs, void* out,
                             uint64_t out_capacity, tdbg_column_view* cols, uint32_t n_cols, uint64_t* n_rows);

uint64_t tdbg_result_num_rows(const tdbg_result* r);
uint32_t tdbg_result_num_columns(const tdbg_result* r);
int32_t tdbg_result_column(const tdbg_result* r, uint32_t i, tdbg_column_view* out);
void tdbg_result_free(tdbg_result* r);

#ifdef __cplusplus
}
#endif

#endif /* THINDB_GPU_H */
