/*
 * The reference's CPU operator path restated with its own data structures
 * and parallel decomposition, for the CPU baseline that bench.py times on the
 * GPU host.  TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline legs and the
 * tests that pin these functions against the oracle); the product never
 * links it.
 *
 * What is kept from native/prism-executor (and prism-substrait's executor):
 *   - HashTable::build (hash_join.rs:87-141): parallel hash of every build row,
 *     a serial scatter into 64 partitions by h & 63, then one map per
 *     partition built in parallel; each map is a SwissTable (the hashbrown
 *     layout: 16-byte control groups, 7-bit tags, triangular probing) keyed by
 *     the hash itself (PassThroughU64, hash_join.rs:23-39), whose value is a
 *     growable row list (Vec<u32>: one heap block per key, first capacity 4).
 *   - probe_matches_chunked (hash_join.rs:286-309): probe batches in parallel,
 *     each batch probed by one thread into two growable u32 arrays.
 *   - hash_aggregate_batches (hash_aggregate.rs:182-339): non-empty batches cut
 *     into chunks of max(len / threads, 1), one map per chunk keyed by the
 *     FxHash group id with std's SipHash-1-3 as the map hash, every aggregate
 *     column cast to a fresh f64 array, per row the accumulators of every
 *     aggregate; chunk maps merged in chunk order; groups in first-appearance
 *     order.
 *   - the executor's Filter -> Project -> Aggregate chain
 *     (executor.rs:91-187): evaluate_predicate + filter_batch per batch in
 *     parallel (every column copied), project_batch_with_exprs per batch in
 *     parallel (BinaryOp casts both sides to f64, a Literal becomes a
 *     full-length array, filter_project.rs:215-294), then the aggregation.
 *   - the fused join -> aggregate (executor.rs:333-458): probe_matches_chunked,
 *     per probe batch the joined (attr, payload) columns by take and a
 *     hash_aggregate into a partial batch, then the merge aggregation.
 * Threads: a pool of `threads` workers taking tasks from a shared counter
 * (rayon's pool runs the same independent tasks).
 */
#ifndef PRISM_REF_CPU_H
#define PRISM_REF_CPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rc_table rc_table;

/* HashTable::build on an Int64 key column (the fast path: hash_i64_fast). */
rc_table* rc_join_build(const int64_t* keys, int64_t n, int32_t threads);
void rc_join_free(rc_table* t);

/* probe_matches_chunked over batches of batch_rows probe rows.  Returns the
 * total pair count; *build_checksum = sum of the build indices of all pairs,
 * *probe_checksum = sum of the (batch-local) probe indices. */
int64_t rc_probe_matches_chunked(const rc_table* t, const int64_t* keys, int64_t n, int64_t batch_rows,
                                 int32_t threads, uint64_t* probe_checksum, uint64_t* build_checksum);

/* Result of an aggregation restated here: groups in output order. */
typedef struct {
  int64_t n_groups;
  uint64_t key_hash[1024];
  int64_t first_row[1024];  /* global row of the group's first row in the aggregated batches */
  double agg[4][1024];      /* GroupAccumulator::result per aggregate */
} rc_groups;

/* hash_aggregate of one 1M-row style batch set: key i64, value i64;
 * SUM, COUNT, MIN, MAX of value grouped by key (at most 1024 groups). */
int rc_groupby_i64(const int64_t* key, const int64_t* value, int64_t n, int64_t batch_rows, int32_t threads,
                   rc_groups* out);

/* Filter(date <= date_max) -> Project([flag_a, flag_b], value) -> Aggregate(
 * group (flag_a, flag_b): SUM, COUNT, AVG of the value), over batches of
 * batch_rows rows; value = price * (100 - discount) (variant 0) or
 * price * 0.01 * (1 - discount * 0.01) (variant 1).  Flags are Utf8 (offsets
 * + bytes).  first_row is the row within the filtered, projected batches,
 * counted globally. */
int rc_filter_project_aggregate(const int32_t* date, const int64_t* price, const int64_t* discount,
                                const uint8_t* fa, const int32_t* fa_off, const uint8_t* fb, const int32_t* fb_off,
                                int64_t n, int32_t date_max, int32_t variant, int64_t batch_rows, int32_t threads,
                                rc_groups* out);

/* Join -> aggregate as the executor's fused path runs it (pairs materialised
 * per probe batch): group the join output by the build attribute, COUNT and
 * SUM of the probe payload.  Writes per-attribute totals into count[attr] and
 * sum[attr] (f64, as the reference outputs them; attr in [0, n_attr)).
 * Returns the number of join pairs processed. */
int64_t rc_join_aggregate(const rc_table* t, const int32_t* build_attr, const int64_t* probe_key,
                          const int64_t* probe_payload, int64_t n_probe, int64_t batch_rows, int32_t threads,
                          int32_t n_attr, double* count, double* sum);

#ifdef __cplusplus
}
#endif

#endif
