/*
 * prism_hip.h -- C ABI of the MI355X (gfx950) hash join / hash aggregation /
 * fused filter-project library.
 *
 * Two layers, both plain C (no torch, no C++ types in any signature):
 *
 *  1. Device layer (prism_ctx_*, prism_join_*, prism_groupby_*, prism_datagen_*):
 *     column buffers already resident in HBM, passed as device pointers.  This
 *     is what a caller that keeps tables in HBM across requests binds (the
 *     operator service, the benchmark).
 *
 *  2. Arrow layer (prism_hash_join*, prism_hash_aggregate*, prism_filter_*,
 *     prism_evaluate_*, prism_project_*): the operator API of
 *     native/prism-executor/src/{hash_join,hash_aggregate,filter_project}.rs
 *     over the Arrow C Data Interface (ArrowArray/ArrowSchema, host memory),
 *     so a Rust `-sys` crate binds it with arrow's `ffi` feature
 *     (native/Cargo.toml:20, prism-jni/src/lib.rs:9).
 *
 * Every entry point returns a prism_status; a non-zero status leaves a
 * thread-local message readable through prism_last_error().  Nothing aborts or
 * unwinds across this boundary.
 */
#ifndef PRISM_HIP_H
#define PRISM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Arrow C Data Interface (ABI fixed by the Apache Arrow specification).      */
/* ------------------------------------------------------------------------- */
#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE
#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4
struct ArrowSchema {
  const char* format;
  const char* name;
  const char* metadata;
  int64_t flags;
  int64_t n_children;
  struct ArrowSchema** children;
  struct ArrowSchema* dictionary;
  void (*release)(struct ArrowSchema*);
  void* private_data;
};
struct ArrowArray {
  int64_t length;
  int64_t null_count;
  int64_t offset;
  int64_t n_buffers;
  int64_t n_children;
  const void** buffers;
  struct ArrowArray** children;
  struct ArrowArray* dictionary;
  void (*release)(struct ArrowArray*);
  void* private_data;
};
#endif /* ARROW_C_DATA_INTERFACE */

/* ------------------------------------------------------------------------- */
/* Status codes: one per PrismError variant (prism-executor/src/error.rs:5-27) */
/* ------------------------------------------------------------------------- */
typedef int32_t prism_status;
#define PRISM_OK 0
#define PRISM_ERR_ARROW 1                   /* PrismError::Arrow */
#define PRISM_ERR_SCHEMA_MISMATCH 2         /* PrismError::SchemaMismatch */
#define PRISM_ERR_INVALID_ARGUMENT 3        /* PrismError::InvalidArgument */
#define PRISM_ERR_EXECUTION 4               /* PrismError::Execution (HIP errors land here) */
#define PRISM_ERR_JOIN 5                    /* PrismError::Join */
#define PRISM_ERR_UNSUPPORTED_AGGREGATION 6 /* PrismError::UnsupportedAggregation */
#define PRISM_ERR_INTERNAL 7                /* PrismError::Internal */
#define PRISM_ERR_RESOURCE 8                /* service: HBM admission refused (no PrismError counterpart) */

/* Message of the last failed call on this thread ("" if none). */
const char* prism_last_error(void);
/* Library version string (replaces Java_io_prism_bridge_PrismNativeExecutor_version,
 * prism-jni/src/lib.rs:249-259). */
const char* prism_version(void);

/* ------------------------------------------------------------------------- */
/* Context: device, HIP stream, scratch memory.                               */
/* ------------------------------------------------------------------------- */
typedef struct prism_ctx prism_ctx;

/* Create a context on HIP device `device` with its own non-blocking stream. */
prism_status prism_ctx_create(int device, prism_ctx** out);
void prism_ctx_destroy(prism_ctx* ctx);
/* Launch subsequent work on an existing hipStream_t (NULL = the context's own). */
prism_status prism_ctx_set_stream(prism_ctx* ctx, void* hip_stream);
void* prism_ctx_stream(prism_ctx* ctx);
prism_status prism_ctx_synchronize(prism_ctx* ctx);
/* Device bytes held by the context and every object created from it. */
uint64_t prism_ctx_device_bytes(prism_ctx* ctx);

/* Device memory and timing for callers without an allocator of their own
 * (tests, benchmarks).  No counterpart in the reference. */
prism_status prism_device_alloc(prism_ctx* ctx, uint64_t bytes, void** out);
prism_status prism_device_free(prism_ctx* ctx, void* p);
/* Synchronous copies on the context stream. */
prism_status prism_copy_to_device(prism_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes);
prism_status prism_copy_to_host(prism_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes);
/* Asynchronous device-to-device copy on the context stream. */
prism_status prism_copy_device(prism_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes);
/* Elapsed device time between begin and end markers recorded on the context
 * stream (HIP events); end synchronises. */
prism_status prism_ctx_timer_begin(prism_ctx* ctx);
prism_status prism_ctx_timer_end(prism_ctx* ctx, double* ms);
/* Per-kernel timing: while enabled, the library brackets each launch of its
 * main kernels with HIP events on the context stream, under these names:
 * "join_probe" (lookup + in-tile compaction), "join_probe_move" (pairs to
 * their final offsets), "groupby_update" (fused filter/aggregate pass).
 * _read synchronises, returns the summed device time and launch count of
 * one name since the last read, and clears that record. */
prism_status prism_ctx_kernel_timing(prism_ctx* ctx, int32_t enable);
prism_status prism_ctx_kernel_timing_read(prism_ctx* ctx, const char* kernel, double* total_ms,
                                          uint64_t* launches);

/* Kernels generated at run time (no counterpart in the reference, whose
 * operators are compiled Rust): group-by update kernels are specialised to
 * their plan and compiled with hipRTC on first use; PRISM_JIT=0 in the
 * environment runs the prebuilt interpreting kernels instead.  Counters of
 * this process; a failed compile falls back to the prebuilt kernel and keeps
 * the compiler log for prism_jit_last_error. */
typedef struct {
  uint64_t kernels;  /* distinct kernels compiled */
  uint64_t failures; /* failed compiles / loads */
  uint64_t hits;     /* requests served by a kernel compiled earlier */
  double compile_ms; /* total compile + load time */
} prism_jit_stats;
prism_status prism_jit_get_stats(prism_jit_stats* out);
const char* prism_jit_last_error(void);

/* ------------------------------------------------------------------------- */
/* Column descriptors (device layer: every pointer is a device pointer).      */
/* ------------------------------------------------------------------------- */
#define PRISM_TYPE_INT32 1
#define PRISM_TYPE_INT64 2
#define PRISM_TYPE_FLOAT64 3
#define PRISM_TYPE_UTF8 4
#define PRISM_TYPE_BOOL 5
#define PRISM_TYPE_DATE32 6

typedef struct {
  int32_t type;            /* PRISM_TYPE_* */
  int32_t reserved;
  int64_t offset;          /* Arrow slice offset in elements (bits for BOOL and validity) */
  const void* data;        /* values; BOOL bit-packed; UTF8 bytes */
  const int32_t* offsets;  /* UTF8 only: i32 offsets (length + 1 entries after `offset`) */
  const uint8_t* validity; /* Arrow validity bitmap, or NULL when every row is valid */
} prism_column;

/* ------------------------------------------------------------------------- */
/* Row programs: the fused predicate / projection stage.                      */
/* A program is postfix code over one row; Predicate (filter_project.rs:22-51) */
/* and ScalarExpr (filter_project.rs:195-213) trees compile to it.            */
/* ------------------------------------------------------------------------- */
#define PRISM_OP_LOAD_I 1      /* push column[arg] (INT32/INT64/DATE32/BOOL) as int64 */
#define PRISM_OP_LOAD_F 2      /* push column[arg] (FLOAT64) */
#define PRISM_OP_CONST_I 3     /* push i64 */
#define PRISM_OP_CONST_F 4     /* push f64 */
#define PRISM_OP_I2F 5         /* top := (double)top */
#define PRISM_OP_ADD_F 6
#define PRISM_OP_SUB_F 7
#define PRISM_OP_MUL_F 8
#define PRISM_OP_DIV_F 9
#define PRISM_OP_NEG_F 10
#define PRISM_OP_ADD_I 11      /* exact int64; overflow raises PRISM_ERR_EXECUTION */
#define PRISM_OP_SUB_I 12
#define PRISM_OP_MUL_I 13
#define PRISM_OP_NEG_I 14
#define PRISM_OP_CMP_I 15      /* pop b, a; push (a <arg> b) on int64; arg = PRISM_CMP_* */
#define PRISM_OP_CMP_F 16      /* same on f64 in IEEE total order (f64::total_cmp, as arrow-ord's
                                  cmp kernels compare floats): NaN == NaN, -0.0 < +0.0 */
#define PRISM_OP_CMP_STR 17    /* push (utf8 column[arg] <i64 >> 8 cmp> string[i64 & 0xff]) */
#define PRISM_OP_LIKE 18       /* push (utf8 column[arg] LIKE pattern string[i64]) */
#define PRISM_OP_ILIKE 19      /* ASCII case-insensitive LIKE */
#define PRISM_OP_IS_NULL 20    /* push column[arg] is null (never null) */
#define PRISM_OP_IS_NOT_NULL 21
#define PRISM_OP_AND 22        /* null if either side null (arrow `and`) */
#define PRISM_OP_OR 23         /* null if either side null (arrow `or`) */
#define PRISM_OP_NOT 24
#define PRISM_OP_CONST_BOOL 25 /* push i64 != 0 */
#define PRISM_OP_SELECT 26     /* pop else, then, cond; push cond ? then : else (zip) */
#define PRISM_OP_CONST_NULL 27 /* push a null */
#define PRISM_OP_ZIP 28        /* pop then, cond, else; push cond ? then : else (else pushed first) */

#define PRISM_CMP_EQ 0
#define PRISM_CMP_NE 1
#define PRISM_CMP_LT 2
#define PRISM_CMP_LE 3
#define PRISM_CMP_GT 4
#define PRISM_CMP_GE 5

typedef struct {
  int32_t op;  /* PRISM_OP_* */
  int32_t arg; /* column index / PRISM_CMP_* */
  int64_t i64;
  double f64;
} prism_instr;

typedef struct {
  const char* data; /* host memory; copied at plan creation */
  uint32_t len;
} prism_string;

/* ------------------------------------------------------------------------- */
/* Hash join, device layer (hash_join.rs:87-200, int64 single-key fast path). */
/* ------------------------------------------------------------------------- */
typedef struct prism_join_table prism_join_table;

#define PRISM_TABLE_AUTO 0   /* pick DENSE when the key range is compact, else HASHED */
#define PRISM_TABLE_HASHED 1 /* open addressing on hash_i64_fast(key), linear probing */
#define PRISM_TABLE_DENSE 2  /* direct-address array over [min_key, max_key] */

/* Probe forms of a table (prism_join_build_opts.probe). */
#define PRISM_PROBE_DIRECT 0       /* every probe row looks its key up in the table in HBM */
#define PRISM_PROBE_PARTITIONED 1  /* radix-partitioned: probe rows are grouped by the hash bits that pick the
                                      table's partition, each partition's slots are loaded into LDS once and
                                      probed there, and the pairs are put back in probe order */

typedef struct {
  uint64_t n_rows;        /* build rows */
  uint64_t n_distinct;    /* distinct keys */
  uint64_t max_dup;       /* largest row list of one key */
  int64_t min_key, max_key;
  int32_t kind;           /* PRISM_TABLE_HASHED / PRISM_TABLE_DENSE */
  int32_t unique;         /* 1 when every key occurs once */
  uint64_t slots;         /* slots (HASHED) or key range (DENSE) */
  uint64_t device_bytes;  /* HBM held by the table */
  uint64_t partition_slots; /* HASHED: slots of one partition (= slots when unpartitioned) */
  uint32_t partitions;    /* HASHED: slots / partition_slots */
  int32_t probe;          /* PRISM_PROBE_* */
} prism_join_table_info;

/* HashTable::build (hash_join.rs:87-141) on a device-resident int64 key column.
 * Rows of one key are kept in ascending row order (as the reference's serial
 * partition pass does, hash_join.rs:116-118).  `kind` = PRISM_TABLE_*. */
prism_status prism_join_build_i64(prism_ctx* ctx, const int64_t* d_keys, uint64_t n_rows,
                                  int32_t kind, prism_join_table** out);
/* The same build with one int32 build column carried in the table entries
 * (the reference's HashTable keeps the build batch, hash_join.rs:135-140): a
 * probe given no d_build_payload sums this column over its matches, reading it
 * with the row id instead of gathering it. */
prism_status prism_join_build_i64_payload(prism_ctx* ctx, const int64_t* d_keys, const int32_t* d_payload,
                                          uint64_t n_rows, int32_t kind, prism_join_table** out);
/* Table layout and probe form, for the build with options (no counterpart in
 * the reference, whose HashTable::build has no knobs; hash_join.rs:41-45 fixes
 * 64 partitions).  A HASHED table of `slots` = next power of two of
 * n / load_factor is cut into slots / partition_slots partitions by the top
 * bits of hash_i64_fast(key); a probe run wraps inside its partition, so a
 * partition of partition_slots x 16 B (4096 slots: 64 KB) is an
 * open-addressing table that fits LDS.  The build refuses a layout where a
 * partition would hold partition_slots keys or more. */
typedef struct {
  int32_t kind;             /* PRISM_TABLE_* (PARTITIONED probes need PRISM_TABLE_HASHED) */
  int32_t probe;            /* PRISM_PROBE_*; PARTITIONED needs unique keys and 256..8192 partition_slots */
  uint64_t partition_slots; /* HASHED: slots per partition, a power of two; 0 = one partition */
  double load_factor;       /* HASHED: keys per slot in (0, 0.9]; 0 = 0.5 */
} prism_join_build_opts;
prism_status prism_join_build_i64_opts(prism_ctx* ctx, const int64_t* d_keys, const int32_t* d_payload,
                                       uint64_t n_rows, const prism_join_build_opts* opts, prism_join_table** out);

/* Autotuner of the table layout and probe form (no counterpart in the
 * reference): builds a table of the keys (and payload) for each candidate --
 * the AUTO table probed directly, HASHED tables of load factor 0.5 / 0.7 probed
 * directly, and the same cut into partitions of 1024..8192 slots and probed
 * partitioned, where the partitions fit the probe's limits -- times
 * one probe of the sample with pairs and payload sums on the context stream,
 * and writes the fastest candidate's options to *best.  cands (room for cap
 * entries, may be NULL) receives every candidate tried, in order. */
typedef struct {
  prism_join_build_opts opts;
  double build_ms;  /* device time of the build */
  double probe_ms;  /* device time of one probe of the sample */
  int32_t status;   /* PRISM_OK, or why the candidate was not built */
  int32_t reserved;
} prism_join_tune_candidate;
prism_status prism_join_tune(prism_ctx* ctx, const int64_t* d_keys, const int32_t* d_payload, uint64_t n_rows,
                             const int64_t* d_probe_sample, const int64_t* d_probe_payload, uint64_t n_sample,
                             prism_join_build_opts* best, prism_join_tune_candidate* cands, uint32_t cap,
                             uint32_t* n_cands);
prism_status prism_join_table_info_get(const prism_join_table* t, prism_join_table_info* out);
void prism_join_table_free(prism_join_table* t);

/* Result of one probe, written to device memory by the probe kernels. */
typedef struct {
  uint64_t matches;         /* total matched (probe, build) pairs, even beyond capacity */
  uint64_t probe_sum_lo;    /* Σ probe_payload over matched pairs, int128 two's complement */
  int64_t probe_sum_hi;
  uint64_t build_sum_lo;    /* Σ build_payload[build_row] over matched pairs, int128 */
  int64_t build_sum_hi;
  uint32_t error;           /* non-zero: a bounded wait gave up (results invalid) */
  uint32_t reserved;
} prism_probe_result;

/* HashTable::probe (hash_join.rs:143-200).  Emits pairs in the reference's
 * order: probe row ascending, build rows ascending within a probe row.
 * d_probe_idx / d_build_idx may both be NULL (count/sum only, the fused
 * join->aggregate form).  Indices are written as row + probe_row_base (probe)
 * and build row.  d_probe_payload (int64) and d_build_payload (int32) are
 * optional; when given, their sums over the matched pairs are accumulated
 * exactly in int128.  Asynchronous on the context stream: `d_result` is device
 * memory and is valid after prism_ctx_synchronize.  When matches > capacity,
 * only `capacity` pairs are written and the caller re-probes with more room. */
prism_status prism_join_probe_i64(prism_ctx* ctx, const prism_join_table* t,
                                  const int64_t* d_probe_keys, uint64_t n_probe,
                                  uint32_t probe_row_base,
                                  const int64_t* d_probe_payload, const int32_t* d_build_payload,
                                  uint32_t* d_probe_idx, uint32_t* d_build_idx, uint64_t capacity,
                                  prism_probe_result* d_result);

/* The same probe without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The table is built
 * beforehand; every output is the caller's.  Its scratch is a fixed 128 MiB
 * device array per GPU, held until the process ends.  A table probed
 * directly takes the one-pass probe here (pairs written at offsets found by
 * looking back over the tiles' match counts, no pair staging): 40 bytes per
 * 4,096 probe rows, so any u32 row count fits.  A partitioned table
 * (PRISM_PROBE_PARTITIONED) stages about 22 bytes per probe row plus 8 bytes
 * per (partition, 4,096-row tile), about 5.9M rows at 256 partitions; a
 * larger probe fails with PRISM_ERR_INVALID_ARGUMENT before anything is
 * launched.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several contexts or streams) race, and must be ordered by the
 * caller.  Results equal prism_join_probe_i64's. */
prism_status prism_join_probe_i64_noalloc(prism_ctx* ctx, const prism_join_table* t,
                                          const int64_t* d_probe_keys, uint64_t n_probe,
                                          uint32_t probe_row_base,
                                          const int64_t* d_probe_payload, const int32_t* d_build_payload,
                                          uint32_t* d_probe_idx, uint32_t* d_build_idx, uint64_t capacity,
                                          prism_probe_result* d_result);

/* ------------------------------------------------------------------------- */
/* Hash aggregation, device layer (hash_aggregate.rs:182-339).                */
/* ------------------------------------------------------------------------- */
#define PRISM_AGG_SUM 0
#define PRISM_AGG_COUNT 1
#define PRISM_AGG_AVG 2
#define PRISM_AGG_MIN 3
#define PRISM_AGG_MAX 4
#define PRISM_AGG_COUNT_DISTINCT 5

typedef struct {
  int32_t func;  /* PRISM_AGG_* */
  int32_t value; /* index of the value program that feeds it */
} prism_groupby_agg;

typedef struct {
  uint32_t n_columns;
  const int32_t* column_types;      /* PRISM_TYPE_* per input column */
  const uint32_t* key_columns;      /* group-by columns (FxHash of their values) */
  uint32_t n_keys;
  const prism_instr* filter;        /* predicate program, or NULL */
  uint32_t n_filter;
  const prism_instr* values;        /* value programs, concatenated */
  const uint32_t* value_lengths;    /* length of each value program */
  uint32_t n_values;
  const prism_groupby_agg* aggs;
  uint32_t n_aggs;
  const prism_string* strings;      /* literals / LIKE patterns used by the programs */
  uint32_t n_strings;
} prism_groupby_plan;

typedef struct prism_groupby prism_groupby;

/* Per-group outputs, host arrays of at least n_groups entries (any may be NULL).
 * Groups come in first-appearance order (hash_aggregate.rs:302-314). */
typedef struct {
  uint64_t* first_row;     /* global row (row_base + row) of each group's first row */
  uint64_t* key_hash;      /* 64-bit FxHash group id (hash_aggregate.rs:483-504) */
  uint64_t* row_count;     /* rows in the group */
  double** agg_f64;        /* [n_aggs]: Float64 result as the reference outputs it */
  int64_t** agg_i64;       /* [n_aggs]: Int64 result for COUNT / COUNT DISTINCT */
  uint64_t** agg_sum128;   /* [n_aggs]: exact int128 SUM (lo, hi pairs) for integer inputs */
} prism_groupby_output;

/* expected_groups: sizing hint (0 = unknown); the table grows as needed. */
prism_status prism_groupby_create(prism_ctx* ctx, const prism_groupby_plan* plan,
                                  uint64_t expected_groups, prism_groupby** out);
/* Stream one batch of rows into the resident group state (asynchronous). */
prism_status prism_groupby_update(prism_groupby* g, const prism_column* columns,
                                  uint64_t n_rows, uint64_t row_base);
/* Finish: order groups by first appearance; returns the group count (syncs). */
prism_status prism_groupby_finish(prism_groupby* g, uint64_t* n_groups);
/* Copy the finished groups into host arrays. */
prism_status prism_groupby_read(prism_groupby* g, const prism_groupby_output* out);
/* Clear the state for reuse without freeing device memory. */
prism_status prism_groupby_reset(prism_groupby* g);
void prism_groupby_destroy(prism_groupby* g);

/* Source of the generated group-by update kernel of `plan` for one launch
 * shape (form 0 = wave pre-aggregation, 1 = per-row slots; bit c of
 * `nullable` = column c carries a validity bitmap), copied NUL-terminated into
 * `buf` (truncated to `cap` bytes); *len = full length.  With `compile` set
 * it is also compiled for gfx950 (no device needed) and a compile error
 * returns PRISM_ERR_EXECUTION with the compiler log as the last error.  A
 * diagnostic with no counterpart in the reference. */
prism_status prism_groupby_jit_source(const prism_groupby_plan* plan, int32_t form, uint32_t nullable,
                                      int32_t compile, char* buf, uint64_t cap, uint64_t* len);


/* ========================================================================= */
/* Arrow layer: prism-executor's operator API over the Arrow C Data          */
/* Interface.  A RecordBatch travels as a struct array ("+s") + its schema.  */
/* Inputs are borrowed (host memory); outputs are owned by the caller, who   */
/* releases them through their `release` callbacks.                          */
/* ========================================================================= */

/* JoinType (hash_join.rs:48-56) */
#define PRISM_JOIN_INNER 0
#define PRISM_JOIN_LEFT 1
#define PRISM_JOIN_RIGHT 2
#define PRISM_JOIN_FULL 3
#define PRISM_JOIN_LEFT_SEMI 4
#define PRISM_JOIN_LEFT_ANTI 5

/* HashJoinConfig (hash_join.rs:58-64): probe = left, build = right. */
typedef struct {
  int32_t join_type;
  const uint32_t* build_keys;
  uint32_t n_build_keys;
  const uint32_t* probe_keys;
  uint32_t n_probe_keys;
} prism_hash_join_config;

/* HashTable (hash_join.rs:77-84): device-resident build side. */
typedef struct prism_hash_table prism_hash_table;

/* HashTable::build (hash_join.rs:87-141) */
prism_status prism_hash_table_build(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                    const uint32_t* key_cols, uint32_t n_keys, prism_hash_table** out);
/* HashTable::probe (hash_join.rs:143-200): two UInt32 arrays, probe rows
 * ascending and build rows ascending within a probe row. */
prism_status prism_hash_table_probe(prism_ctx* ctx, prism_hash_table* table, const struct ArrowArray* probe,
                                    const struct ArrowSchema* probe_schema, const uint32_t* probe_keys,
                                    uint32_t n_keys, struct ArrowArray* out_probe_idx,
                                    struct ArrowSchema* out_probe_schema, struct ArrowArray* out_build_idx,
                                    struct ArrowSchema* out_build_schema);
void prism_hash_table_free(prism_hash_table* table);

/* hash_join (hash_join.rs:312-365): output = probe columns, then build columns
 * (a name clash gets "_right", hash_join.rs:438-449). */
prism_status prism_hash_join(prism_ctx* ctx, const struct ArrowArray* probe, const struct ArrowSchema* probe_schema,
                             const struct ArrowArray* build, const struct ArrowSchema* build_schema,
                             const prism_hash_join_config* cfg, struct ArrowArray* out,
                             struct ArrowSchema* out_schema);

/* probe_matches_chunked (hash_join.rs:286-309): one (probe_idx, build_idx) pair
 * of UInt32 arrays per probe batch (batch-local probe indices).  out_idx_schema
 * receives the UInt32 schema once. */
prism_status prism_probe_matches_chunked(prism_ctx* ctx, const struct ArrowArray* const* probe_batches,
                                         uint32_t n_batches, const struct ArrowSchema* probe_schema,
                                         const struct ArrowArray* build, const struct ArrowSchema* build_schema,
                                         const uint32_t* probe_keys, const uint32_t* build_keys, uint32_t n_keys,
                                         struct ArrowArray* out_probe_idx, struct ArrowArray* out_build_idx,
                                         struct ArrowSchema* out_idx_schema);

/* hash_join_probe_chunked (hash_join.rs:206-282): joined batches, empty ones
 * dropped; Right/Full run as per-batch Inner as in the reference.  out_batches
 * has room for n_batches arrays; *n_out tells how many were written. */
prism_status prism_hash_join_probe_chunked(prism_ctx* ctx, const struct ArrowArray* const* probe_batches,
                                           uint32_t n_batches, const struct ArrowSchema* probe_schema,
                                           const struct ArrowArray* build, const struct ArrowSchema* build_schema,
                                           const prism_hash_join_config* cfg, struct ArrowArray* out_batches,
                                           struct ArrowSchema* out_schema, uint32_t* n_out);

/* AggExpr / HashAggConfig (hash_aggregate.rs:88-101) */
typedef struct {
  uint32_t column;
  int32_t func;            /* PRISM_AGG_* */
  const char* output_name;
} prism_agg_expr;

#define PRISM_AGG_FLAG_EXACT_DECIMAL 1 /* SUM of integer columns as Decimal128(38,0), exact */

typedef struct {
  const uint32_t* group_by;
  uint32_t n_group_by;
  const prism_agg_expr* aggregates;
  uint32_t n_aggregates;
  uint32_t flags;
} prism_hash_agg_config;

/* hash_aggregate (hash_aggregate.rs:174-176) */
prism_status prism_hash_aggregate(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                  const prism_hash_agg_config* cfg, struct ArrowArray* out,
                                  struct ArrowSchema* out_schema);
/* hash_aggregate_batches (hash_aggregate.rs:182-339, global path 343-447) */
prism_status prism_hash_aggregate_batches(prism_ctx* ctx, const struct ArrowArray* const* batches, uint32_t n_batches,
                                          const struct ArrowSchema* schema, const prism_hash_agg_config* cfg,
                                          struct ArrowArray* out, struct ArrowSchema* out_schema);

/* ------------------------------------------------------------------------- */
/* Distributed aggregate (prism-bench handler.rs:270-344, 397-526): each     */
/* producer runs the plan with the expanded aggregate list and ships its     */
/* partial batch (Flight stays on the host); the reducer merges them.        */
/* ------------------------------------------------------------------------- */
/* expand_worker_aggregates (handler.rs:397-428): AVG becomes SUM + COUNT of
 * the same column; worker aggregate i is named "agg_<i>".  out_worker has
 * room for 2 * n entries; out_map[2i] (and out_map[2i+1] for an AVG, else
 * UINT32_MAX) are the worker indices of original aggregate i.  The names
 * stay valid until the next call on the same thread. */
prism_status prism_expand_worker_aggregates(const prism_agg_expr* original, uint32_t n, prism_agg_expr* out_worker,
                                            uint32_t* n_worker, uint32_t* out_map);
/* Reducer merge of partial batches (group columns first, then the worker
 * aggregates of `original` in expansion order): build_final_merge_config
 * (handler.rs:430-457: SUM for SUM / COUNT / COUNT DISTINCT partials, MIN,
 * MAX), hash_aggregate_batches on the device, then
 * reconstruct_original_aggregate_batch (handler.rs:459-526): COUNT and COUNT
 * DISTINCT as Int64 (`as i64` of the merged sum), AVG as Float64 sum / count
 * (0.0 when the count is 0), SUM / MIN / MAX as Float64; nulls read as 0. */
prism_status prism_merge_partial_aggregates(prism_ctx* ctx, const struct ArrowArray* const* partials,
                                            uint32_t n_partials, const struct ArrowSchema* schema,
                                            uint32_t group_by_count, const prism_agg_expr* original,
                                            uint32_t n_original, struct ArrowArray* out,
                                            struct ArrowSchema* out_schema);

/* ------------------------------------------------------------------------- */
/* Arrow IPC stream ingest: prism-jni's executePlan receives its input as an  */
/* Arrow IPC stream in a JNI byte array and decodes it with arrow's           */
/* StreamReader (prism-jni/src/lib.rs:96-121).  Streams of Int32 / Int64 /    */
/* Float64 / Utf8 / Bool / Date32 fields; dictionaries, body compression and  */
/* nested types are rejected with PRISM_ERR_ARROW.                            */
/* ------------------------------------------------------------------------- */
typedef struct prism_ipc_stream prism_ipc_stream; /* a decoded stream resident in HBM */

/* The stream's record batches as Arrow C Data struct arrays whose buffers
 * point into `bytes` (no copy: `bytes` must outlive the arrays), with one
 * schema; out_batches has room for `capacity` arrays.  No device needed. */
prism_status prism_ipc_read_host(const uint8_t* bytes, uint64_t n_bytes, struct ArrowArray* out_batches,
                                 uint32_t capacity, uint32_t* n_batches, struct ArrowSchema* out_schema);
/* The stream decoded straight into HBM: every column's buffers copied once
 * from `bytes` through pinned staging; the columns feed the device-level
 * entries (group-by update, join build/probe, service requests). */
prism_status prism_ipc_read_device(prism_ctx* ctx, const uint8_t* bytes, uint64_t n_bytes, prism_ipc_stream** out);
prism_status prism_ipc_stream_info(const prism_ipc_stream* s, uint32_t* n_batches, uint32_t* n_columns);
/* Field `column`: name (valid while the stream lives), PRISM_TYPE_*, nullability. */
prism_status prism_ipc_field(const prism_ipc_stream* s, uint32_t column, const char** name, int32_t* type,
                             int32_t* nullable);
/* Rows of record batch `batch` and the device descriptor of its column `column`. */
prism_status prism_ipc_column(const prism_ipc_stream* s, uint32_t batch, uint32_t column, uint64_t* n_rows,
                              prism_column* out);
void prism_ipc_stream_free(prism_ipc_stream* s);

/* Arrow IPC stream writer of prism-jni's reply (prism-jni/src/lib.rs:111-121,
 * arrow's StreamWriter over the plan's output batches): the schema message,
 * one record batch message per batch (struct arrays sharing `schema`, sliced
 * inputs written from their offset), then the end-of-stream marker, into a
 * host buffer the library allocates (free it with prism_ipc_buffer_free).
 * With no batches the output is empty, as the reference writes no stream
 * when the plan produced none.  Types: the operator types (Int32, Int64,
 * UInt32, Float64, Utf8, Boolean, Date32, Decimal128(38,0)).  No device needed. */
prism_status prism_ipc_write(const struct ArrowArray* batches, uint32_t n_batches, const struct ArrowSchema* schema,
                             uint8_t** out_bytes, uint64_t* out_len);
void prism_ipc_buffer_free(uint8_t* bytes);

/* ------------------------------------------------------------------------- */
/* Parquet decode: prism-executor's parquet_scan reads each file's footer,    */
/* skips row groups whose statistics rule the predicate out and decodes the  */
/* projected column chunks with the `parquet` crate (parquet_scan.rs:55-258, */
/* 167-258, 357-494).  Here the footer and statistics come from           */
/* prism_parquet_open / prism_parquet_statistics (the caller keeps           */
/* row_group_might_match), and the chosen column chunks are decoded into     */
/* HBM.  Flat schemas of BOOLEAN, INT32 (int32, date), INT64, DOUBLE and      */
/* BYTE_ARRAY (utf8) leaves; PLAIN, dictionary (PLAIN_DICTIONARY /           */
/* RLE_DICTIONARY) and RLE boolean pages, v1 and v2; codecs UNCOMPRESSED,    */
/* SNAPPY, GZIP.  Anything else is rejected with PRISM_ERR_ARROW.            */
/* ------------------------------------------------------------------------- */
typedef struct prism_parquet_file prism_parquet_file;   /* parsed footer over the caller's bytes */
typedef struct prism_parquet_table prism_parquet_table; /* decoded column chunks resident in HBM */

typedef struct {
  int32_t present;       /* the column chunk carries Statistics */
  int32_t physical_type; /* Parquet physical type: 0 BOOLEAN, 1 INT32, 2 INT64, 5 DOUBLE, 6 BYTE_ARRAY, ... */
  int64_t null_count;    /* -1 when absent */
  int32_t has_min, has_max;
  const uint8_t* min;    /* PLAIN encoding of the value (little-endian numbers, raw bytes) */
  uint32_t min_len;
  const uint8_t* max;
  uint32_t max_len;
} prism_parquet_stats;

/* Parse the footer of a whole Parquet file held in `bytes` (which must
 * outlive the handle).  Host only. */
prism_status prism_parquet_open(const uint8_t* bytes, uint64_t n_bytes, prism_parquet_file** out);
prism_status prism_parquet_file_info(const prism_parquet_file* f, uint32_t* n_row_groups, uint32_t* n_columns,
                                     int64_t* n_rows);
/* Leaf `column`: name, PRISM_TYPE_* (0 when the leaf cannot be decoded), nullability, physical type. */
prism_status prism_parquet_field(const prism_parquet_file* f, uint32_t column, const char** name, int32_t* type,
                                 int32_t* nullable, int32_t* physical_type);
prism_status prism_parquet_row_group(const prism_parquet_file* f, uint32_t row_group, int64_t* n_rows);
/* Column chunk statistics (row-group skipping, parquet_scan.rs:357-494); pointers live with the handle. */
prism_status prism_parquet_statistics(const prism_parquet_file* f, uint32_t row_group, uint32_t column,
                                      prism_parquet_stats* out);
void prism_parquet_close(prism_parquet_file* f);
/* Decode row groups x columns into HBM (NULL lists mean all), one batch per row group. */
prism_status prism_parquet_read_device(prism_ctx* ctx, const prism_parquet_file* f, const uint32_t* row_groups,
                                       uint32_t n_row_groups, const uint32_t* columns, uint32_t n_columns,
                                       prism_parquet_table** out);
prism_status prism_parquet_table_info(const prism_parquet_table* t, uint32_t* n_batches, uint32_t* n_columns);
prism_status prism_parquet_table_column(const prism_parquet_table* t, uint32_t batch, uint32_t column,
                                        uint64_t* n_rows, prism_column* out);
/* Batch `batch` copied to the host as an Arrow C Data struct array (the RecordBatch parquet_scan returns). */
prism_status prism_parquet_table_export(const prism_parquet_table* t, uint32_t batch, struct ArrowArray* out,
                                        struct ArrowSchema* out_schema);
void prism_parquet_table_free(prism_parquet_table* t);

/* ------------------------------------------------------------------------- */
/* Join -> aggregate by a build-side attribute without materialising pairs:  */
/* the fused join-aggregate of the prism-substrait executor                  */
/* (executor.rs:333-458) for                                                 */
/*   SELECT b.attr, COUNT(*), SUM(p.payload) FROM probe p JOIN build b        */
/*   ON p.key = b.key GROUP BY b.attr                                         */
/* Exact (int128 sums, u64 counts).  Groups come out in first appearance     */
/* over the join output (probe row ascending, then build row ascending).     */
/* Device pointers; keys and attributes carry no nulls.  The build arrays    */
/* must stay valid until finish.                                             */
/* ------------------------------------------------------------------------- */
typedef struct prism_join_agg prism_join_agg;
prism_status prism_join_agg_create(prism_ctx* ctx, const int64_t* d_build_keys, const int32_t* d_build_attr,
                                   uint64_t n_build, prism_join_agg** out);
/* Accumulate one probe batch; probe_row_base = global row of its first row.
 * d_probe_payload may be NULL (sums stay 0). */
prism_status prism_join_agg_probe(prism_join_agg* j, const int64_t* d_probe_keys, const int64_t* d_probe_payload,
                                  uint64_t n_probe, uint64_t probe_row_base);
prism_status prism_join_agg_finish(prism_join_agg* j, uint64_t* n_groups);
/* Host arrays of n_groups entries (any may be NULL).  first_pair = (probe row
 * << 32) | build row of the group's first join output row. */
prism_status prism_join_agg_read(prism_join_agg* j, int32_t* attr, uint64_t* count, uint64_t* sum_lo,
                                 int64_t* sum_hi, uint64_t* first_pair);
prism_status prism_join_agg_reset(prism_join_agg* j);
void prism_join_agg_destroy(prism_join_agg* j);

/* ------------------------------------------------------------------------- */
/* Operator service (no counterpart in the reference; the north star's       */
/* device-resident service for prism-jni): continuous batching of group-by,  */
/* probe and filter requests into persistent tile-queue kernels, resident    */
/* build tables with HBM admission and LRU eviction, graph replay of         */
/* recurring batch shapes, an online tile/occupancy tuner, p50/p99 latency.  */
/* Request inputs are device-resident; they must stay valid until the        */
/* request completes.  The dispatcher thread polls for submissions for 5 ms  */
/* after its last work before it sleeps (one busy host core while requests   */
/* keep coming; PRISM_SVC_IDLE_SPIN_US sets the window in microseconds, 0    */
/* sleeps at once).                                                          */
/* ------------------------------------------------------------------------- */
typedef struct prism_service prism_service;
typedef struct {
  uint64_t hbm_budget;          /* bytes for resident tables + working arenas (0: 64 GiB) */
  uint32_t max_inflight;        /* batches on the device at once, one stream each (0: 3, max 8) */
  uint32_t max_batch_requests;  /* requests per batch (0 or > 256: 256) */
  uint64_t max_batch_rows;      /* rows per batch (0: 256 Mi) */
  uint64_t arena_bytes;         /* working HBM per in-flight batch (0: 256 MiB) */
  int32_t use_graphs;           /* replay captured launch sequences for recurring batch shapes: 0 never, 1 batches
                                   of 10 or more kernel launches (a graph launch costs the host about as much as
                                   ten kernel launches), 2 every recurring shape */
  int32_t autotune;             /* pick tile rows / workgroups per CU per kind from the first batches */
} prism_service_config;

#define PRISM_REQ_GROUPBY 1     /* target = group-by plan id */
#define PRISM_REQ_PROBE 2       /* target = table id: matches + int128 payload sums (+ pairs) */
#define PRISM_REQ_FILTER 3      /* target = filter plan id: selected rows (+ optional bitmap) */
typedef struct {
  int32_t kind;
  uint32_t target;
  const prism_column* columns;    /* GROUPBY / FILTER: device columns, as many as the plan declares */
  const int64_t* probe_keys;      /* PROBE */
  const int64_t* probe_payload;   /* PROBE, may be NULL */
  uint64_t n_rows;
  unsigned long long* out_mask;   /* FILTER: ceil(n_rows / 64) device words, may be NULL */
  /* PROBE, optional: HashTable::probe's output (hash_join.rs:143-200), the
   * matched (probe row, build row) pairs in probe order, as
   * prism_join_probe_i64 writes them; at most pair_capacity pairs are written
   * (the result's count holds every match).  NULL = counts and sums only. */
  uint32_t* probe_idx;
  uint32_t* build_idx;
  uint64_t pair_capacity;
} prism_request;

typedef struct {
  int32_t status;       /* prism_status of the request */
  int32_t kind;
  uint64_t n_rows;
  uint64_t count;       /* GROUPBY: groups; PROBE: matches; FILTER: selected rows */
  uint64_t probe_sum_lo;
  int64_t probe_sum_hi;
  uint64_t build_sum_lo;
  int64_t build_sum_hi;
  double latency_ms;    /* submit -> result on the host */
  double queue_ms;      /* submit -> admitted into a batch */
  uint64_t batch_id;
} prism_request_result;

typedef struct {
  uint64_t submitted, completed, batches, tables_built, tables_evicted, resident_table_bytes, graph_launches;
  double p50_ms, p99_ms, mean_batch_requests;
  double mean_batch_gpu_ms, max_batch_gpu_ms;  /* stream time from batch start to its results on the host */
  double mean_batch_launch_ms;                 /* host time to admit and launch a batch */
  double mean_batch_detect_ms;                 /* host wall time after the launch returned beyond the
                                                  batch's device time (completion detection), >= 0 */
  double mean_batch_rows;
  double mean_batch_kernel_ms;                 /* batch start -> its kernels done (before the results copy) */
  uint32_t tile_rows[3];      /* per kind: group-by, probe, filter */
  uint32_t blocks_per_cu[3];
  int32_t tuned[3];
} prism_service_stats;

prism_status prism_service_create(prism_ctx* ctx, const prism_service_config* cfg, prism_service** out);
void prism_service_destroy(prism_service* s);
/* Resident build table with unique keys (host arrays are copied; the table is
 * rebuilt from them after an eviction). */
prism_status prism_service_register_table(prism_service* s, uint32_t table_id, const int64_t* h_keys,
                                          const int32_t* h_payload, uint64_t n);
/* Group-by plan (same plan as prism_groupby_create) holding at most max_groups
 * (1..1024) groups per request. */
/* Resident build tables under store keys, as prism-flight's PartitionStore
 * keeps batches under string keys (shuffle_writer.rs:42-86; prism-bench names
 * a table by its `store_key`, handler.rs:800-843).  put: append the rows to the
 * key's table, creating it (with a fresh id >= 2^31, written to *table_id) on
 * the first put; the table is rebuilt and its keys must stay unique (a batch
 * that repeats a key is refused and the key keeps its earlier rows).
 * table_id: the id of a key's table, for prism_request.target.  clear: drop a
 * key's table (absent keys are a no-op).  drop_exchange: drop every table
 * whose key starts with "exchange/<exchange_id>/" (PartitionStore::drop_exchange).
 * A table that requests wait for or run on is not changed: the call fails and
 * the caller retries after they complete. */
prism_status prism_service_put_table(prism_service* s, const char* key, const int64_t* h_keys, const int32_t* h_payload,
                                     uint64_t n, uint32_t* table_id);
prism_status prism_service_table_id(prism_service* s, const char* key, uint32_t* table_id);
prism_status prism_service_clear_table(prism_service* s, const char* key);
prism_status prism_service_drop_exchange(prism_service* s, const char* exchange_id);
prism_status prism_service_register_groupby(prism_service* s, uint32_t plan_id, const prism_groupby_plan* plan,
                                            uint64_t max_groups);
/* Filter plan: a boolean row program (the compiled form of a Predicate). */
prism_status prism_service_register_filter(prism_service* s, uint32_t plan_id, const prism_instr* code, uint32_t n_code,
                                           const prism_string* strings, uint32_t n_strings, const int32_t* column_types,
                                           uint32_t n_columns);
prism_status prism_service_submit(prism_service* s, const prism_request* req, uint64_t* ticket);
prism_status prism_service_poll(prism_service* s, uint64_t ticket, int32_t* done);
prism_status prism_service_wait(prism_service* s, uint64_t ticket, prism_request_result* result);
/* Groups of a finished GROUPBY request, as prism_groupby_read writes them. */
prism_status prism_service_read_groups(prism_service* s, uint64_t ticket, const prism_groupby_output* out);
prism_status prism_service_release(prism_service* s, uint64_t ticket);
prism_status prism_service_drain(prism_service* s);
prism_status prism_service_stats_get(prism_service* s, prism_service_stats* out);
prism_status prism_service_reset_stats(prism_service* s);

/* ------------------------------------------------------------------------- */
/* Arrow layer: filter / project (filter_project.rs:20-333) and LIKE          */
/* (string_ops.rs:16-25).  Predicate and ScalarExpr trees travel as node     */
/* arrays; a node refers to its children by index into the same array.      */
/* ------------------------------------------------------------------------- */
/* ScalarValue (filter_project.rs:54-62) */
#define PRISM_SCALAR_INT32 0
#define PRISM_SCALAR_INT64 1
#define PRISM_SCALAR_FLOAT64 2
#define PRISM_SCALAR_UTF8 3
#define PRISM_SCALAR_BOOLEAN 4
#define PRISM_SCALAR_DATE32 5
typedef struct {
  int32_t type;     /* PRISM_SCALAR_* */
  int32_t reserved;
  int64_t i64;      /* Int32 / Int64 / Boolean / Date32 */
  double f64;       /* Float64 */
  const char* str;  /* Utf8 bytes */
  uint32_t str_len;
  uint32_t reserved2;
} prism_scalar;

/* Predicate (filter_project.rs:21-51) */
#define PRISM_PRED_LITERAL 0
#define PRISM_PRED_EQ 1
#define PRISM_PRED_NE 2
#define PRISM_PRED_LT 3
#define PRISM_PRED_LE 4
#define PRISM_PRED_GT 5
#define PRISM_PRED_GE 6
#define PRISM_PRED_IS_NULL 7
#define PRISM_PRED_IS_NOT_NULL 8
#define PRISM_PRED_AND 9
#define PRISM_PRED_OR 10
#define PRISM_PRED_NOT 11
#define PRISM_PRED_LIKE 12
#define PRISM_PRED_ILIKE 13
typedef struct {
  int32_t kind;          /* PRISM_PRED_* */
  uint32_t column;       /* comparisons, IS [NOT] NULL, LIKE */
  int32_t left, right;   /* AND / OR children; NOT uses left */
  prism_scalar value;    /* comparisons: the literal; LITERAL: i64 != 0 */
  const char* pattern;   /* LIKE / ILIKE */
  uint32_t pattern_len;
  uint32_t reserved;
} prism_pred_node;

/* ScalarExpr (filter_project.rs:185-213) */
#define PRISM_EXPR_COLUMN 0
#define PRISM_EXPR_LITERAL 1
#define PRISM_EXPR_BINARY 2
#define PRISM_EXPR_NEGATE 3
#define PRISM_EXPR_IF_THEN 4
/* ArithmeticOp (filter_project.rs:185-192) */
#define PRISM_ARITH_ADD 0
#define PRISM_ARITH_SUB 1
#define PRISM_ARITH_MUL 2
#define PRISM_ARITH_DIV 3
typedef struct {
  int32_t kind;              /* PRISM_EXPR_* */
  uint32_t column;           /* COLUMN */
  prism_scalar value;        /* LITERAL */
  int32_t op;                /* BINARY: PRISM_ARITH_* */
  int32_t left, right;       /* BINARY children; NEGATE uses left (expression indices) */
  const int32_t* clause_preds; /* IF_THEN: predicate roots (indices into the predicate nodes) */
  const int32_t* clause_exprs; /* IF_THEN: THEN expressions (indices into the expression nodes) */
  uint32_t n_clauses;
  int32_t else_expr;         /* IF_THEN: ELSE expression index, or -1 */
} prism_expr_node;

/* evaluate_predicate (filter_project.rs:64-123): BooleanArray mask.  nodes[0]
 * is the root. */
prism_status prism_evaluate_predicate(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                      const prism_pred_node* nodes, uint32_t n_nodes, struct ArrowArray* out_mask,
                                      struct ArrowSchema* out_schema);
/* filter_batch (filter_project.rs:125-135): rows whose mask entry is true
 * (null = false), every column. */
prism_status prism_filter_batch(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                const struct ArrowArray* mask, const struct ArrowSchema* mask_schema,
                                struct ArrowArray* out, struct ArrowSchema* out_schema);
/* project_batch (filter_project.rs:137-160): the selected columns in order;
 * n_indices = 0 gives a batch with no columns and the input row count. */
prism_status prism_project_batch(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                 const uint32_t* indices, uint32_t n_indices, struct ArrowArray* out,
                                 struct ArrowSchema* out_schema);
/* filter_and_project (filter_project.rs:162-180): nodes may be NULL (no
 * predicate); n_indices = 0 keeps every column.  Only the projected columns
 * are compacted on the device. */
prism_status prism_filter_and_project(prism_ctx* ctx, const struct ArrowArray* batch, const struct ArrowSchema* schema,
                                      const prism_pred_node* nodes, uint32_t n_nodes, const uint32_t* indices,
                                      uint32_t n_indices, struct ArrowArray* out, struct ArrowSchema* out_schema);
/* evaluate_scalar_expr (filter_project.rs:215-294): expression root `root`. */
prism_status prism_evaluate_scalar_expr(prism_ctx* ctx, const struct ArrowArray* batch,
                                        const struct ArrowSchema* schema, const prism_expr_node* exprs,
                                        uint32_t n_exprs, const prism_pred_node* preds, uint32_t n_preds,
                                        int32_t root, struct ArrowArray* out, struct ArrowSchema* out_schema);
/* project_batch_with_exprs (filter_project.rs:304-333): passthrough columns,
 * then one column "expr_{i}" per expression root. */
prism_status prism_project_batch_with_exprs(prism_ctx* ctx, const struct ArrowArray* batch,
                                            const struct ArrowSchema* schema, const uint32_t* indices,
                                            uint32_t n_indices, const prism_expr_node* exprs, uint32_t n_exprs,
                                            const int32_t* roots, uint32_t n_roots, const prism_pred_node* preds,
                                            uint32_t n_preds, struct ArrowArray* out, struct ArrowSchema* out_schema);
/* Device-level predicate: a boolean row program over device columns (the
 * compiled form of a Predicate).  Writes ceil(n_rows / 64) mask words (bit set =
 * row selected: value true and not null) when d_mask is not NULL, and the
 * number of selected rows.  Synchronous. */
prism_status prism_filter_mask(prism_ctx* ctx, const prism_instr* code, uint32_t n_code, const prism_string* strings,
                               uint32_t n_strings, const int32_t* column_types, const prism_column* columns,
                               uint32_t n_columns, uint64_t n_rows, unsigned long long* d_mask, uint64_t* selected);
/* string_like / string_ilike (string_ops.rs:16-25) over a Utf8 array. */
prism_status prism_string_like(prism_ctx* ctx, const struct ArrowArray* array, const struct ArrowSchema* schema,
                               const char* pattern, uint32_t pattern_len, int32_t case_insensitive,
                               struct ArrowArray* out, struct ArrowSchema* out_schema);

/* ------------------------------------------------------------------------- */
/* Synthetic inputs (seeded, counter-based; bit-identical to the numpy twin    */
/* in prism_executor.datagen).  Device pointers.                              */
/* ------------------------------------------------------------------------- */
/* out[i] = lo + floor(u32(i) * range / 2^32), u32(i) = high word of rnd(seed, stream, i) */
prism_status prism_datagen_uniform_i64(prism_ctx* ctx, int64_t* d_out, uint64_t n, uint64_t seed,
                                       uint32_t stream, int64_t lo, uint64_t range);
prism_status prism_datagen_uniform_i32(prism_ctx* ctx, int32_t* d_out, uint64_t n, uint64_t seed,
                                       uint32_t stream, int64_t lo, uint64_t range);
/* out[i] = (int64) rnd(seed, stream, i): full 64-bit range */
prism_status prism_datagen_bits_i64(prism_ctx* ctx, int64_t* d_out, uint64_t n, uint64_t seed,
                                    uint32_t stream);
/* out[i] = perm(i): a seeded bijection of [0, n) (Feistel with cycle walking) */
prism_status prism_datagen_permutation_i64(prism_ctx* ctx, int64_t* d_out, uint64_t n,
                                           uint64_t seed, uint32_t stream);
/* out[i] = perm_V(rank - 1), rank ~ discrete CDF d_cdf[V] (u64 fixed point) */
prism_status prism_datagen_zipf_i64(prism_ctx* ctx, int64_t* d_out, uint64_t n, uint64_t seed,
                                    uint32_t stream, const uint64_t* d_cdf, uint64_t n_values,
                                    uint32_t perm_stream);
/* Utf8 column of one-byte strings: bytes[i] = alphabet[uniform(0, n_alpha)], offsets[i] = i */
prism_status prism_datagen_flag_utf8(prism_ctx* ctx, uint8_t* d_bytes, int32_t* d_offsets,
                                     uint64_t n, uint64_t seed, uint32_t stream,
                                     const char* alphabet, uint32_t n_alpha);

/* Random lowercase Utf8 rows: length_i = min_len + uniform(max_len - min_len + 1)
 * (stream), byte p of the data = 'a' + uniform(26) (stream + 1).  Writes n + 1
 * offsets and *total_bytes bytes (at most bytes_capacity). */
prism_status prism_datagen_lower_utf8(prism_ctx* ctx, uint8_t* d_bytes, uint64_t bytes_capacity, int32_t* d_offsets,
                                      uint64_t n, uint64_t seed, uint32_t stream, uint32_t min_len, uint32_t max_len,
                                      uint64_t* total_bytes);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PRISM_HIP_H */
