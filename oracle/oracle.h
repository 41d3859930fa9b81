/*
 * oracle.h — CPU restatement of thinDB's grouped-aggregate semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU engine is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (libthindb_gpu.so, thindb_amd/) never links,
 * imports or calls anything here.
 *
 * The plan format is the C-ABI's `tdbg_plan_desc` (include/thindb_gpu.h):
 * the oracle reads the same request the GPU operator receives.
 */
#ifndef THINDB_ORACLE_H
#define THINDB_ORACLE_H

#include <stdint.h>
#include "../include/thindb_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One column of a host table: `n_rows` values of `type` (enum tdbg_type),
 * optional validity bitmap (bit 1 = valid, bit r at byte r>>3, bit r&7). */
typedef struct or_column {
  uint32_t type;
  uint32_t pad;
  const void* data;
  const uint8_t* validity;
} or_column;

/* A table as contiguous columns, cut into row groups of `rg_rows` rows
 * (the unit workers claim, src/exec/rg_tiles.zig:1-30). */
typedef struct or_table {
  uint64_t n_rows;
  uint32_t rg_rows;
  uint32_t n_cols;
  const or_column* cols;
} or_table;

typedef struct or_result or_result;

/* Runs the grouped query with `dop` workers.  dop == 1 is the reference's
 * deterministic single-worker order.  Returns 0, TDBG_ERR_UNSUPPORTED for
 * shapes the reference's low-card/silo cores decline, or TDBG_ERR_INVALID. */
int32_t oracle_group_by(const tdbg_plan_desc* plan, const or_table* t, uint32_t dop, or_result** out);
/* The same result through the silo's high-cardinality arrangement (rows
 * routed by key hash into buckets, buckets split into parts, each part
 * folded into its own table): the CPU baseline for high-cardinality
 * queries.  COUNT(DISTINCT) returns TDBG_ERR_UNSUPPORTED. */
int32_t oracle_group_by_radix(const tdbg_plan_desc* plan, const or_table* t, uint32_t dop, or_result** out);
uint64_t oracle_result_num_rows(const or_result* r);
uint32_t oracle_result_num_columns(const or_result* r);
int32_t oracle_result_column(const or_result* r, uint32_t i, tdbg_column_view* out);
/* For a DOUBLE SUM/AVG output column: the compensated (Neumaier) value of
 * each emitted row, the accuracy reference for the 1e-12 tolerance. */
int32_t oracle_result_compensated(const or_result* r, uint32_t i, double* out);
/* Number of groups before OFFSET/LIMIT. */
uint64_t oracle_result_num_groups(const or_result* r);
void oracle_result_free(or_result* r);
const char* oracle_last_error(void);
/* Tests only: plants one known mistake in the ORDER BY comparator (0 = none),
 * so a test can show that its checks catch that mistake. */
void oracle_plant(int which);

#ifdef __cplusplus
}
#endif

#endif
