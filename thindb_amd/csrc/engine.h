// engine.h — structures shared by the host runtime and the gfx950 kernels of
// libthindb_gpu.so.  Everything here is plain data passed by value as kernel
// arguments (uniform across a launch, read through scalar loads).
#pragma once

#if defined(__HIPCC_RTC__)
// Compiled at run time by hipRTC (plan-specialised kernels): the runtime
// headers are implicit and the fixed-width types live in __hip_internal.
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace tdbg {

constexpr int kMaxCols = 16;    // distinct columns one plan reads
constexpr int kMaxPreds = 8;    // merged predicate terms
constexpr int kMaxParts = 8;    // GROUP BY key parts (src/exec/v2_lowcard_group.zig:114)
constexpr int kMaxAggs = 16;    // aggregates (src/exec/v2_lowcard_group.zig:115)
constexpr int kVec = 4;         // consecutive rows per thread per strip
constexpr uint32_t kRowAlign = 256;   // row-group starts in an arena are multiples of this
constexpr uint64_t kEmptyKey = ~0ull; // empty hash slot (relative keys stay below 2^63)
constexpr int kLdsProbe = 32;   // LDS probe limit before a row spills to HBM
constexpr uint32_t kLdsBudget = 160u * 1024u - 1024u;

// Physical element kinds of a staged column.
enum ColKind : uint32_t { K_U8 = 0, K_I8, K_I16, K_I32, K_I64, K_F32, K_F64 };

struct ColDesc {
  const void* data;        // arena base; element r is physical row r
  const uint32_t* valid;   // validity bitmap words (bit 1 = valid) or null
  uint32_t kind;
  uint32_t pad;
};

// A predicate term after merging the WHERE's leaves per column:
// integer ranges are inclusive [lo_i, hi_i]; float ranges carry flags.
enum PredKind : uint32_t { P_RANGE_I = 0, P_RANGE_F, P_NEQ_I, P_NEQ_F };
struct PredDesc {
  uint32_t kind;
  uint32_t col;
  uint32_t lo_incl, hi_incl;  // float ranges
  int64_t lo_i, hi_i;
  double lo_f, hi_f;
};

// Key part: relative value rel = v - base, 0 <= rel < span, folded into a
// mixed-radix group index key = sum(rel_i * stride_i).  A nullable part keeps
// rel 0 for NULL and its base one below the smallest value, so NULL sorts
// first.  A packed 128-bit key (G_WIDE) holds rel as a bit field instead:
// bits [bit_off, bit_off + bit_width) of the key, parts from the first at the
// low end, so the last part is the most significant as in the mixed radix;
// there span 0 stands for 2^64 (a full-range BIGINT).
struct PartDesc {
  uint32_t col;
  uint32_t nullable;
  int64_t base;
  uint64_t span;
  uint64_t stride;
  uint32_t bit_off;
  uint32_t bit_width;  // 0: mixed radix (G_WIDE: a constant part, rel 0)
};

// A group key as stored in records: the relative key in `lo` (G_DENSE: the
// record index; G_HASH: the mixed-radix key), or the packed 128-bit key
// (G_WIDE).
struct GKey {
  uint64_t lo, hi;
};

// Aggregate kernels.  Integer SUM/AVG accumulate exactly: in LDS either one
// i64 (when rows-per-workgroup * max|v| < 2^63) or split into a u64 sum of
// the low 32 bits and an i64 sum of the high 32 bits; in HBM as i128.
enum AggKind : uint32_t {
  A_COUNT_COL = 0,  // COUNT(col): non-null count
  A_SUM_I,          // SUM / AVG of integers
  A_SUM_F,          // SUM / AVG of floats (f64)
  A_MIN_I, A_MAX_I, // integer MIN/MAX (i64)
  A_MIN_F, A_MAX_F  // float MIN/MAX through the order-preserving i64 map
};
struct AggDesc {
  uint32_t kind;
  uint32_t col;       // index into cols[]
  uint32_t nullable;  // keeps its own non-null count
  uint32_t split;     // A_SUM_I: LDS keeps split 32-bit halves
  uint32_t lds_off;   // byte offset of this agg's LDS array(s)
  uint32_t lds_ns_off;// byte offset of its LDS non-null counter array
  uint32_t g_off;     // byte offset of its state in the HBM record
  uint32_t g_ns_off;  // byte offset of its non-null count in the record
  uint32_t lo_only;   // A_SUM_I: only the state's low 64 bits are read (a wrapped BIGINT SUM), so the
                      // record holds an 8-byte state and HBM updates are single wrapping 64-bit adds
  uint32_t pad;
};

enum LdsMode : uint32_t { LDS_NONE = 0, LDS_DIRECT = 1, LDS_HASH = 2 };
// G_DENSE: record index = relative key; G_HASH: records carry the u64
// relative key at g_key_off; G_WIDE: records carry a packed 128-bit key (two
// u64 words, low first) at g_key_off.
enum GlobalMode : uint32_t { G_DENSE = 1, G_HASH = 2, G_WIDE = 3 };

struct Tile {
  uint64_t row0;
  uint32_t n;
  uint32_t pad;
};

// Error bits raised by kernels (host maps them to status codes).
enum ErrBit : uint32_t { ERR_KEY_RANGE = 1u, ERR_HASH_FULL = 2u };

// One node of a WHERE tree in postfix order (kinds of HavingKind: H_CMP,
// H_IS_NULL, H_IS_NOT_NULL, H_NOT, H_AND, H_OR).  A comparison is one
// predicate term over column pd.col; IS [NOT] NULL reads pd.col's validity.
constexpr int kMaxFilter = 12;
struct FilterNode {
  uint32_t kind;
  uint32_t arity;  // H_AND / H_OR
  PredDesc pd;
};

struct ScanParams {
  ColDesc cols[kMaxCols];
  PredDesc preds[kMaxPreds];
  FilterNode filt[kMaxFilter];  // AND-ed with preds: the WHERE beyond a conjunction of leaves
  uint32_t n_filt;
  uint32_t pad_filt;
  PartDesc parts[kMaxParts];
  AggDesc aggs[kMaxAggs];
  uint32_t n_cols, n_preds, n_parts, n_aggs;
  // LDS table
  uint32_t lds_mode;
  uint32_t lds_cap;        // slots
  uint32_t lds_parts;      // LDS_DIRECT: workgroups per tile range, each owning
                           // lds_cap consecutive keys (grid % parts == 0 and
                           // (grid / parts) % 8 == 0 keeps partners on one XCD)
  uint32_t lds_key64;      // hash keys are u64 (else u32)
  uint32_t lds_keys_off;   // byte offsets inside dynamic LDS
  uint32_t lds_count_off;
  uint32_t lds_bytes;
  // HBM table of records
  uint32_t g_mode;
  uint32_t g_rec;          // record stride in bytes (multiple of 8)
  uint32_t g_count_off;    // byte offset of the u64 row count in a record
  uint32_t g_key_off;      // byte offset of the u64 key (hash mode)
  uint64_t g_cap;          // records (dense: key span; hash: power of two)
  uint8_t* g_state;
  // work
  const Tile* tiles;
  uint32_t n_tiles;
  uint32_t pad0;
  uint32_t* err;
  uint64_t* g_inserted;    // hash mode: number of claimed records
};

// ---- radix partitioning (high cardinality) ----------------------------------

// A surviving row as the partition passes move it: the relative key (u32;
// hash partitions: the key's 32-bit hash), a validity word when an aggregate
// input is nullable, each kernel aggregate's input value at its column's
// native width, and with hash partitions the key itself last, as key_words
// u64 words at key_off.
struct RowLayout {
  uint32_t stride;             // bytes per row (multiple of 8)
  uint32_t flags_off;          // ~0u: no validity word
  uint32_t key_off;            // hash partitions: byte offset of the key words (multiple of 8)
  uint32_t key_words;          // hash partitions: 1 (G_HASH relative key) or 2 (G_WIDE packed key)
  uint32_t val_off[kMaxAggs];  // byte offset of aggregate a's value
  uint32_t val_bytes[kMaxAggs];
};

struct HavingLeaf;

// Bins of the radix leaves' group-count histogram (ORDER BY COUNT(*) cuts):
// a group lands in bin min(row count, kRxCountBins - 1).
constexpr uint32_t kRxCountBins = 256;
constexpr uint32_t kRxTotalSlots = 64;  // group_total counters, 64 bytes apart

struct RadixParams {
  RowLayout row;
  uint32_t shift1, bits1;      // level-1 digit = key >> shift1 (bits1 wide)
  uint32_t shift2, bits2;      // level-2 digit = (key >> shift2) & (2^bits2 - 1); bits2 = 0: one level
  uint32_t leaf_bits;          // a leaf holds 2^leaf_bits consecutive keys
  uint32_t n_leaves;
  uint32_t tile_rows;          // rows a scatter workgroup stages in LDS at once (multiple of 256)
  uint32_t dense_out;          // leaves write every key of their range at record index key
                               // (G_DENSE table) instead of compacted keyed records
  uint32_t chunks2;            // level-2 workgroups per level-1 bucket (<= 64)
  uint32_t tile_rows2;         // rows a level-2 scatter workgroup stages at once (multiple of 256, <= 8192)
  uint32_t* hist2;             // [2^bits1 * chunks2 * 2^bits2] chunk digit counts, then offsets
  uint32_t* counts_out;        // compacted leaves: each group's row count by record index, or null
  uint32_t* count_hist;        // [kRxCountBins] groups by min(row count, kRxCountBins - 1) (zeroed)
  uint8_t* rows1;              // level-1 partitioned rows
  uint32_t* digits1;           // two levels: each level-1 row's u32 route value (its first word), packed, so
                               // the level-2 count reads 4 bytes per row instead of the row (null: one level)
  uint8_t* rows2;              // level-2 partitioned rows (== rows1 when one level)
  uint64_t* bucket_start;      // [2^bits1 + 1] level-1 bucket starts (from the histogram)
  uint64_t* bucket_fill;       // [2^bits1] level-1 cursors: rows written per bucket
  uint64_t* leaf_start;        // [n_leaves + 1] leaf row ranges in the final array
  uint32_t* leaf_groups;       // [n_leaves] groups each leaf wrote (compacted records from its first row),
                               // then scratch for their positions: [1 + n_leaves / 256 + 1] words
  const HavingLeaf* having;    // compacted leaves: only groups passing this HAVING are written (null: all)
  // ORDER BY COUNT(*) LIMIT over compacted leaves: each leaf writes only its
  // leaf_topk best groups by (row count in topk_desc order, key), a superset
  // of the leaf's share of the query's top leaf_topk (0: every group); every
  // occupied group is counted in group_total[(leaf % kRxTotalSlots) * 8]
  uint32_t leaf_topk;
  uint32_t topk_desc;
  unsigned long long* group_total;
  // Hash partitions (wide or sparse keys): digits are taken from the top of
  // a 32-bit hash of the key (leaf_bits = the hash bits below both digits),
  // and each leaf folds its rows into an LDS hash table of ScanParams.lds_cap
  // slots: u32 tags at tag_off (0 = free), the keys (row.key_words u64 words
  // per slot) at ScanParams.lds_keys_off, the list of claimed slots at
  // list_off, the aggregate arrays as usual.
  uint32_t hashed;
  uint32_t tag_off;
  uint32_t list_off;           // u32 [lds_cap]: the slots a leaf claimed, in claim order
  uint32_t tag_mask;           // tag bits kept (all but in tests, which narrow them to force tag collisions)
  uint32_t hash_exact;         // leaves fold only in the exact (round-based) form
};

// ---- ordering / emission -------------------------------------------------

// One ORDER BY component value source (64-bit order-preserving words).
enum CompSrc : uint32_t {
  C_KEY = 0,      // key part `idx`: i64 value
  C_COUNT,        // row count
  C_U64_AT,       // u64 at record offset (COUNT(col) non-null count)
  C_I64_AT,       // i64 at record offset (MIN/MAX int, wrapped BIGINT SUM low word)
  C_I128_HI,      // high word of an i128 at record offset (DECIMAL SUM)
  C_I128_LO,      // low word (unsigned)
  C_F64_AT,       // f64 at record offset
  C_FORD_AT,      // float MIN/MAX stored through the order map
  C_AVG_I,        // f64(i128 sum) / n / 10^aux  (the emitted AVG)
  C_AVG_F,        // f64 sum / n
  C_AGG_NULL,     // 0 for a group with no non-null input (the aggregate is NULL), else 1
  C_PKEY_VALID,   // key part `idx` of a nullable part: 0 NULL, 1 valid
  C_PKEY          // key part `idx`: the value's two's-complement bits masked to aux bits (0 for NULL)
};
struct CompDesc {
  uint32_t src;
  uint32_t idx;      // key part index
  uint32_t off;      // record offset of the value
  uint32_t ns_off;   // record offset of the non-null count, or ~0 (use count)
  uint32_t desc;
  uint32_t aux;      // C_AVG_I: decimal scale; C_PKEY: the key type's width in bits
};

constexpr int kMaxComps = 24;
struct OrderParams {
  CompDesc comps[kMaxComps];
  uint32_t n_comps;
  uint32_t n_parts;
  PartDesc parts[kMaxParts];
  const uint8_t* g_state;
  uint32_t g_rec, g_count_off, g_key_off, g_mode;
};

// Output column writer for the finalize kernel.
enum OutKind : uint32_t {
  O_KEY = 0,     // key part value, narrowed to the key type
  O_COUNT,       // COUNT(*)
  O_NS,          // COUNT(col)
  O_SUM_I64,     // integer SUM -> BIGINT (wrap)
  O_SUM_I128,    // DECIMAL SUM -> DECIMAL128
  O_SUM_F64,     // float SUM -> DOUBLE
  O_AVG_I,       // integer/decimal AVG -> DOUBLE
  O_AVG_F,       // float AVG -> DOUBLE
  O_MINMAX_I,    // integer MIN/MAX -> input type
  O_MINMAX_F     // float MIN/MAX -> input type
};
struct OutDesc {
  uint32_t kind;
  uint32_t idx;      // key part
  uint32_t bytes;    // element width of the output type
  uint32_t is_f32;   // float output narrowed to f32
  uint32_t off;      // record offset of the state
  uint32_t ns_off;   // record offset of the non-null count or ~0
  uint64_t data_off; // byte offset of the column in the output buffer
  uint64_t valid_off;// byte offset of its validity bytes, or ~0
  double scale_div;  // 10^scale for AVG of DECIMAL
};
constexpr int kMaxOut = kMaxParts + kMaxAggs;
struct FinalParams {
  OutDesc outs[kMaxOut];
  uint32_t n_outs;
  uint32_t n_parts;
  PartDesc parts[kMaxParts];
  const uint8_t* g_state;
  uint32_t g_rec, g_count_off, g_key_off, g_mode;
};

// HAVING node (the C-ABI's tdbg_having_node, postfix order): CMP pushes
// `output[target] op literal` (NULL output -> false), IS_NULL/IS_NOT_NULL
// push the output's nullness, NOT negates the top, AND/OR fold `arity`
// entries (src/exec/v2_shape_group_topn.zig:2502-2535).
enum HavingKind : uint32_t { H_CMP = 0, H_IS_NULL, H_IS_NOT_NULL, H_NOT, H_AND, H_OR };
constexpr uint32_t kMaxHaving = 32;
struct HavingNode {
  uint32_t kind;
  uint32_t target;
  uint32_t op;
  uint32_t arity;
  uint32_t lit_float;
  uint32_t pad;
  int64_t lit_lo, lit_hi;
  double lit_f64;
};

// HAVING evaluated by the radix leaf pass on its LDS table before records
// are written: where each output column's value lives there.
struct HavingSrc {
  uint32_t kind;     // OutKind of the output column
  uint32_t agg;      // ScanParams aggregate holding its state (aggregate outputs)
  uint32_t part;     // key part (O_KEY)
  uint32_t pad;
  double scale_div;  // 10^scale for AVG of DECIMAL
};
struct HavingLeaf {
  uint32_t n;
  uint32_t pad;
  HavingNode nodes[kMaxHaving];
  HavingSrc src[kMaxParts + kMaxAggs];
};

// COUNT(DISTINCT x) of one plan: every filtered row with a non-NULL x adds
// its (group, x) composite, key * span + (x - base), to an open-addressing
// set; a composite's first insertion counts one for its group (`counts`,
// indexed by the group's relative key).
struct DistinctParams {
  ColDesc x;                // the input column
  int64_t base;             // its zone-map minimum
  uint64_t span;            // max - min + 1
  unsigned long long* set;  // 2^set_bits slots, empty = ~0
  uint64_t set_mask;
  uint32_t* counts;         // one per relative key of the plan's domain (not to_records)
  uint32_t dslot_off;       // record offset of the group's distinct count
  uint32_t to_records;      // the plan's HBM records are addressable by key (not radix):
                            // composites use the record index, counts go straight to the records
  // WIDE: a composite that does not fit 63 bits (a float input, or a value span
  // too wide): each slot holds the row index whose entry (group, value bits)
  // it stands for; entries[2r], entries[2r + 1] belong to row r
  uint32_t wide;
  uint32_t is_float;        // WIDE: the value's canonical f64 bits (types.canonicalFloatBits)
  uint32_t bitmap;          // composites index a bit array (`set` as 32-bit words), not hash slots
  uint64_t* entries;
};

// ---- batched execution (one persistent launch over many plans) ------------

// A work item of the persistent kernel: tiles [a0, a1) of plan q for a scan
// or a radix partition pass, leaves [a0, a1) for a radix leaf pass, or
// records [a0, a1) whose groups an emit item writes to the host.
enum BatchItemKind : uint32_t { BI_SCAN = 0, BI_RX_HIST, BI_RX_SCATTER, BI_RX_AGG, BI_EMIT };
struct BatchItem {
  uint32_t kind;
  uint32_t q;
  uint32_t a0, a1;
  uint32_t part;   // BI_SCAN over a partitioned direct table: the key range owned;
                   // BI_RX_HIST / BI_RX_SCATTER / BI_EMIT: the item's index among the plan's
  uint32_t pad;
};

// Ready queues the persistent kernel fills as plans finish a stage (the
// items of the next stage are appended by the workgroup that completed the
// stage's last item).  Deeper stages are taken first.
enum BatchQueue : uint32_t { BQ_SCATTER = 0, BQ_AGG, BQ_EMIT, kBatchQueues };

constexpr int kTmplWords = 64;  // record template words (records up to 512 B)
struct BatchQuery {
  ScanParams sp;
  RadixParams rx;
  OrderParams op;
  FinalParams fp;                // data/valid offsets are absolute in the output buffer
  uint64_t tmpl[kTmplWords];     // one initialised record
  uint32_t n_hist, n_scatter;    // radix partition items of this plan
  uint32_t n_comp;               // items whose completion releases the emit items (scans, or leaves)
  uint32_t init_records;         // records start from tmpl (else the leaf pass writes them all)
  uint32_t sc0, ag0, em0;        // first scatter, leaf and emit item (indices into the item array)
  uint32_t n_emit;               // emit items (records in kEmitChunk blocks)
  uint32_t n_having;             // HAVING nodes at `having` (0 = none)
  uint32_t jit_sig;              // scan plans: its generated strip in a run-time compiled batch kernel (~0u: none)
  const HavingNode* having;
  uint32_t* sync;                // [0] histogram items done, [1] scatter items done, [2] n_comp items done
  uint32_t* item_hist;           // radix: [n_hist][2^bits1] per-item digit counts, then offsets
  uint64_t* emit_status;         // per emit item: kEmitReady | groups it writes, once counted
  uint64_t* n_out;               // groups emitted (output buffer header)
};
constexpr uint64_t kEmitReady = 1ull << 63;

// The persistent kernel's queue state (zeroed before every launch): every
// word on its own 128-byte line, so claims, appends and polls of different
// words do not queue behind each other.
struct alignas(128) BatchWord {
  uint32_t v;
  uint32_t pad[31];
};
struct BatchSched {
  BatchWord head0;               // static queue: items [0, n_static) taken in order
  BatchWord emitters;            // emit items running
  BatchWord head[kBatchQueues];  // ready queues: items taken
  BatchWord tail[kBatchQueues];  // ready queues: items appended (slots reserved)
};

struct BatchArgs {
  const BatchQuery* Q;
  const BatchItem* items;
  uint32_t n_static;             // items [0, n_static) are ready at launch (histograms, scans)
  uint32_t q_items[kBatchQueues];  // items appended to each ready queue over the launch
  BatchSched* sched;
  uint32_t* slots[kBatchQueues]; // per ready queue: item index + 1 per slot (0 = being written)
  uint32_t* err;
  uint8_t* out;                  // the output buffer's base (host memory mapped for the device)
  uint32_t spin_limit;
  uint32_t emit_cap;             // emit items running at once beyond which others are preferred
  uint32_t lds_bytes;            // dynamic LDS per workgroup (emit items stage records in it)
};

// kBatchChunk consecutive records of plan q starting at r0 (record init).
constexpr uint32_t kBatchChunk = 4096;
struct BatchBlock {
  uint32_t q;
  uint32_t pad;
  uint64_t r0;
};

constexpr uint32_t kBatchThreads = 512;
constexpr uint32_t kEmitChunk = 8192;     // records per emit item
constexpr uint32_t kEmitPer = kEmitChunk / kBatchThreads;  // records per thread of an emit item
constexpr uint32_t kEmitStageOff = 2048;  // LDS of an emit item: counters and masks, then staged records
// LDS of the radix scatter: digit tables (counts, offsets, cursors, scan
// temporaries, bucket bases for up to 1024 digits) then the staged rows.
constexpr uint32_t kRxStageOff = 24 * 1024;
enum BatchErr : uint32_t { ERR_BATCH_WAIT = 4u };  // a dependency wait gave up (ErrBit space)

#if !defined(__HIPCC_RTC__)
}  // namespace tdbg
#include <string>
#include <vector>
namespace tdbg {
// ---- kernel launchers (defined in the .hip files) -------------------------
hipError_t launch_fill_records(uint8_t* state, uint64_t n_rec, uint32_t rec_bytes, const uint64_t* tmpl,
                               hipStream_t s, uint32_t* flags = nullptr, uint32_t n_flags = 0);
hipError_t launch_scan_agg(const ScanParams& p, uint32_t grid, uint32_t block, hipStream_t s);
// Plan-specialised kernel (jit.cpp); hipErrorNotSupported = use launch_scan_agg.
// `cached_only`: launch only a kernel an earlier call compiled (compiling
// and loading a module allocates), else hipErrorNotSupported.
hipError_t launch_scan_jit(const ScanParams& p, uint32_t grid, uint32_t block, hipStream_t s, bool* used,
                           bool cached_only = false);
const char* jit_last_log();
// A batch kernel compiled at run time with a generated strip per structure
// in `sigs` (scan plans pick theirs by BatchQuery::jit_sig; `generic`: some
// scan plan has none); nullptr: use launch_batch.
hipFunction_t batch_jit_kernel(const std::vector<ScanParams>& sigs, bool generic);
std::string jit_signature(const ScanParams& p);  // the structure a generated strip fixes
hipError_t launch_batch_jit(hipFunction_t fn, const BatchQuery* Q, const BatchItem* items, const BatchArgs* args,
                            uint32_t grid, uint32_t lds_bytes, hipStream_t s);
hipError_t launch_compact(const OrderParams& p, uint64_t n_rec, uint32_t* block_counts, uint32_t* gids,
                          uint64_t* n_out, hipStream_t s);
hipError_t launch_sort(const OrderParams& p, uint32_t* gids, uint64_t n, uint64_t need, void* scratch,
                       size_t scratch_bytes, uint32_t* out_sel, hipStream_t s);
// The first `need` of m candidate groups gids[idx[0..m)] (any order, m <=
// kSortCandidatesMax, need <= m) in ORDER BY order into out_sel; ties
// break by position in gids.  scratch: sort_candidates_scratch_bytes(m).
constexpr uint32_t kSortCandidatesMax = 4096;
constexpr uint32_t kLeafTopMax = 1024;  // ORDER BY COUNT(*): OFFSET + LIMIT up to which each radix leaf writes only its top
size_t sort_candidates_scratch_bytes(uint32_t m, uint32_t n_comps);
hipError_t launch_sort_candidates(const OrderParams& p, const uint32_t* gids, uint64_t n, const uint32_t* idx,
                                  uint32_t m, uint64_t need, void* scratch, uint32_t* out_sel, hipStream_t s);
size_t sort_scratch_bytes(uint64_t n, uint32_t n_comps);
// HAVING over the compacted positions gids_in[0, *n_in): the survivors, in
// order, to gids_out and their count to *n_kept.  `max_n` bounds *n_in;
// `scratch` holds having_scratch_bytes(max_n) bytes (pass masks, counts).
size_t having_scratch_bytes(uint64_t max_n);
hipError_t launch_having(const FinalParams& p, const HavingNode* nodes, uint32_t n_nodes, const uint32_t* gids_in,
                         const uint64_t* n_in, uint64_t max_n, void* scratch, uint32_t* gids_out, uint64_t* n_kept,
                         hipStream_t s);
// COUNT(DISTINCT): inserts the scan's (group, x) composites (the plan's
// tiles and predicates), then writes each compacted group's count into its
// record at D.dslot_off.  The set must hold at least twice the rows scanned.
constexpr uint32_t kDistinctLdsKeys = 4096;  // domains up to this many groups count in LDS
hipError_t launch_distinct_insert(const ScanParams& sp, const DistinctParams& D, uint64_t n_keys, uint32_t grid,
                                  hipStream_t s);  // n_keys: the relative key domain
hipError_t launch_distinct_gather(const FinalParams& fp, const DistinctParams& D, const uint32_t* gids,
                                  const uint64_t* n_groups, uint64_t max_n, hipStream_t s);
hipError_t launch_finalize(const FinalParams& p, const uint32_t* sel, uint64_t n_out, uint8_t* out,
                           hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

// One workgroup finishes a dense group table of at most kFastRec records:
// occupancy, ORDER BY (radix select of the first varying order word, then a
// bitonic sort of the candidates by every word and the slot) or slot order
// without one, OFFSET/LIMIT, and the typed result columns - in one launch,
// with the group count and the scan's error flags in a header, so the host
// synchronises once per query.  `fallback` in the header: more than
// kFastCand groups tie at the cut (the host runs the general path over the
// same records).
struct FastHeader {
  uint64_t n_groups;  // occupied records
  uint64_t n_out;     // result rows written
  uint32_t err;       // the scan's error flags
  uint32_t fallback;
  uint64_t pad[5];
};
constexpr uint32_t kFastRec = 12288;   // records one workgroup finishes (12 per thread)
constexpr uint32_t kFastNeed = 1024;   // OFFSET + LIMIT it serves
constexpr uint32_t kFastCand = 1024;   // candidates it sorts
constexpr uint32_t kFastWords = 4;     // order words it handles
constexpr uint32_t kFastUnordered = 4096;  // records of an unordered, unlimited result
hipError_t launch_dense_topn(const OrderParams& op, const FinalParams& fp, uint32_t n_rec, uint32_t offset,
                             uint32_t need, const uint32_t* err, uint8_t* out, FastHeader* hdr, hipStream_t s);
hipError_t launch_iota(uint32_t* p, uint64_t n, hipStream_t s);
// Radix partitioning phases (radix.hip): RX_HIST .. RX_AGG in order.
enum RadixPhase : int { RX_HIST = 0, RX_OFFSETS, RX_SCATTER1, RX_SCATTER2, RX_AGG };
hipError_t launch_radix(const ScanParams& p, const RadixParams& r, uint32_t grid, unsigned long long* hist,
                        hipStream_t s, int phase);
// Encoded blocks (decode.hip): out[i] = base + codes[i] (codes `width` bytes,
// the add wrapping at `elem` bytes); out[i] = values[run of row i] for runs
// starting at starts[0..n_runs), rows from `covered` on 0.
hipError_t launch_for_expand(const void* codes, uint32_t width, int64_t base, uint32_t elem, uint64_t n, void* out,
                             hipStream_t s);
hipError_t launch_rle_expand(const void* values, const uint32_t* starts, uint32_t n_runs, uint32_t covered,
                             uint32_t elem, uint64_t n, void* out, hipStream_t s);
// ORDER BY COUNT(*) over compacted leaves (r.counts_out set): the indexes
// into the group positions of every group whose count is >= cut (desc) or
// <= cut (asc), in any order, to idx (cap entries); their number to *n_idx
// (zeroed here).
hipError_t launch_radix_count_take(const RadixParams& r, uint32_t cut, bool desc, uint32_t* idx, uint32_t cap,
                                   uint32_t* n_idx, hipStream_t s);
// After RX_AGG of compacted leaves: the group count to *n_out and the
// groups' record positions, in key order, to gids.
hipError_t launch_radix_gids(const RadixParams& r, uint32_t* gids, uint64_t* n_out, hipStream_t s);
// Batched execution (batch.hip): record init over `blocks`, the persistent
// kernel over `items` (grid workgroups of kBatchThreads, lds_bytes dynamic
// LDS each; *head must start at 0), compaction and finalize.
hipError_t launch_batch_init(const BatchQuery* q, const BatchBlock* blocks, uint32_t n_blocks, hipStream_t s);
// `args` lives in device memory (written before every launch); the kernel
// reads its fields where they are used instead of holding them in registers.
hipError_t launch_batch(const BatchQuery* Q, const BatchItem* items, const BatchArgs* args, uint32_t n_items,
                        uint32_t grid, uint32_t lds_bytes, hipStream_t s);
#endif

}  // namespace tdbg
