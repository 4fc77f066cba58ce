/*
 * CPU restatement of skn-dd/prism-query's hash join / hash aggregation /
 * LIKE operators, used ONLY as the checker in tests/, the parity check in
 * __graft_entry__.smoke() and the CPU-baseline leg of bench.py.  Nothing in the
 * product path links or calls it.
 *
 * Each function cites the reference lines it follows
 * (native/prism-executor/src/...).
 */
#ifndef PRISM_ORACLE_H
#define PRISM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Column types use the same codes as include/prism_hip.h (PRISM_TYPE_*). */
#define ORC_INT32 1
#define ORC_INT64 2
#define ORC_FLOAT64 3
#define ORC_UTF8 4
#define ORC_BOOL 5
#define ORC_DATE32 6

typedef struct {
  int32_t type;
  int32_t reserved;
  int64_t length;
  const void* data;       /* values (BOOL: one byte per row, 0/1) */
  const int32_t* offsets; /* UTF8 only */
  const uint8_t* valid;   /* one byte per row (1 = valid), or NULL */
} orc_column;

typedef struct {
  int64_t n_rows;
  int32_t n_columns;
  int32_t reserved;
  const orc_column* columns;
} orc_batch;

/* AggFunc (hash_aggregate.rs:79-86) */
#define ORC_SUM 0
#define ORC_COUNT 1
#define ORC_AVG 2
#define ORC_MIN 3
#define ORC_MAX 4
#define ORC_COUNT_DISTINCT 5

typedef struct {
  int32_t column;
  int32_t func;
} orc_agg;

/* Result of orc_hash_aggregate.  Arrays are sized by the caller to at least
 * the total row count (the largest possible number of groups). */
typedef struct {
  int64_t n_groups;
  int64_t* first_batch; /* batch of each group's first row (key source, hash_aggregate.rs:302-314) */
  int64_t* first_row;   /* row in that batch */
  uint64_t* key_hash;   /* FxHash group id */
  double* f64;          /* [n_aggs][cap]: GroupAccumulator::result (hash_aggregate.rs:161-170) */
  int64_t* i64;         /* [n_aggs][cap]: COUNT / COUNT DISTINCT as Int64 (hash_aggregate.rs:326-329) */
  uint64_t* sum_lo;     /* [n_aggs][cap]: exact int128 SUM of the integer inputs (not in the reference) */
  int64_t* sum_hi;
  int64_t cap;
} orc_agg_out;

/* FxHash of one row's group keys (hash_aggregate.rs:24-75,483-504). */
uint64_t orc_group_hash(const orc_batch* b, const int32_t* key_cols, int32_t n_keys, int64_t row);

/* hash_aggregate_batches (hash_aggregate.rs:182-339) incl. the global path
 * (343-447).  n_threads plays rayon::current_num_threads() (chunking, 217).
 * Returns 0, or -1 on an unsupported input. */
int orc_hash_aggregate(const orc_batch* batches, int32_t n_batches, const int32_t* group_by, int32_t n_gb,
                       const orc_agg* aggs, int32_t n_aggs, int32_t n_threads, orc_agg_out* out);

/* HashTable::build + HashTable::probe (hash_join.rs:87-200): matched
 * (probe_row, build_row) pairs in the reference's emission order.  Keys are the
 * listed columns (single Int64 key = fast path without key check).  Returns the
 * pair count; writes at most `cap` pairs. */
int64_t orc_join_probe(const orc_batch* build, const int32_t* build_keys, const orc_batch* probe,
                       const int32_t* probe_keys, int32_t n_keys, uint32_t* out_probe, uint32_t* out_build,
                       int64_t cap);

/* SQL LIKE / ILIKE as arrow's like kernels evaluate it (string_ops.rs:16-25). */
int orc_like(const uint8_t* s, int32_t ns, const uint8_t* p, int32_t np, int32_t case_insensitive);
void orc_like_column(const orc_column* c, const uint8_t* p, int32_t np, int32_t case_insensitive, uint8_t* out);

/* Eager join -> aggregate by a build column (executor.rs:333-458 semantics,
 * aggregated without materialising the pairs):
 *   for every build row b: count[attr[b]] += cnt_probe[key[b]],
 *                          sum[attr[b]]   += sum_probe[key[b]]  (int128).
 * attr values must lie in [0, n_groups).  Returns 0. */
int orc_join_agg_eager(const int64_t* build_key, const int32_t* build_attr, int64_t n_build,
                       const int64_t* probe_key, const int64_t* probe_payload, int64_t n_probe,
                       int64_t n_groups, uint64_t* count_out, uint64_t* sum_lo_out, int64_t* sum_hi_out);

#ifdef __cplusplus
}
#endif

#endif
