"""ctypes binding of libprism_hip.so (the C ABI declared in include/prism_hip.h).

The library is required: every entry point here raises when it is missing, and
there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_int, c_int32, c_int64,
                    c_uint32, c_uint64, c_void_p)

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG_DIR), "lib", "libprism_hip.so")

# Status codes (one per PrismError variant, native/prism-executor/src/error.rs:5-27).
PRISM_OK = 0
STATUS_NAMES = {
    1: "Arrow",
    2: "SchemaMismatch",
    3: "InvalidArgument",
    4: "Execution",
    5: "Join",
    6: "UnsupportedAggregation",
    7: "Internal",
    8: "Resource",
}

TYPE_INT32, TYPE_INT64, TYPE_FLOAT64, TYPE_UTF8, TYPE_BOOL, TYPE_DATE32 = 1, 2, 3, 4, 5, 6

OP = dict(LOAD_I=1, LOAD_F=2, CONST_I=3, CONST_F=4, I2F=5, ADD_F=6, SUB_F=7, MUL_F=8, DIV_F=9,
          NEG_F=10, ADD_I=11, SUB_I=12, MUL_I=13, NEG_I=14, CMP_I=15, CMP_F=16, CMP_STR=17,
          LIKE=18, ILIKE=19, IS_NULL=20, IS_NOT_NULL=21, AND=22, OR=23, NOT=24, CONST_BOOL=25,
          SELECT=26, CONST_NULL=27, ZIP=28)
CMP = dict(EQ=0, NE=1, LT=2, LE=3, GT=4, GE=5)
AGG = dict(SUM=0, COUNT=1, AVG=2, MIN=3, MAX=4, COUNT_DISTINCT=5)
TABLE_AUTO, TABLE_HASHED, TABLE_DENSE = 0, 1, 2


class PrismError(Exception):
    """Mirror of PrismError: `.kind` is the variant name, `.status` the C code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status
        self.kind = STATUS_NAMES.get(status, "Internal")


class Column(Structure):
    _fields_ = [("type", c_int32), ("reserved", c_int32), ("offset", c_int64), ("data", c_void_p),
                ("offsets", c_void_p), ("validity", c_void_p)]


class Instr(Structure):
    _fields_ = [("op", c_int32), ("arg", c_int32), ("i64", c_int64), ("f64", c_double)]


class PString(Structure):
    _fields_ = [("data", c_char_p), ("len", c_uint32)]


class JoinTableInfo(Structure):
    _fields_ = [("n_rows", c_uint64), ("n_distinct", c_uint64), ("max_dup", c_uint64),
                ("min_key", c_int64), ("max_key", c_int64), ("kind", c_int32), ("unique", c_int32),
                ("slots", c_uint64), ("device_bytes", c_uint64), ("partition_slots", c_uint64),
                ("partitions", c_uint32), ("probe", c_int32)]


PROBE_DIRECT, PROBE_PARTITIONED = 0, 1


class BuildOpts(Structure):
    _fields_ = [("kind", c_int32), ("probe", c_int32), ("partition_slots", c_uint64), ("load_factor", c_double)]


class TuneCandidate(Structure):
    _fields_ = [("opts", BuildOpts), ("build_ms", c_double), ("probe_ms", c_double), ("status", c_int32),
                ("reserved", c_int32)]


class ProbeResult(Structure):
    _fields_ = [("matches", c_uint64), ("probe_sum_lo", c_uint64), ("probe_sum_hi", c_int64),
                ("build_sum_lo", c_uint64), ("build_sum_hi", c_int64), ("error", c_uint32),
                ("reserved", c_uint32)]

    @property
    def probe_sum(self) -> int:
        return (self.probe_sum_hi << 64) + self.probe_sum_lo

    @property
    def build_sum(self) -> int:
        return (self.build_sum_hi << 64) + self.build_sum_lo


class GroupbyAgg(Structure):
    _fields_ = [("func", c_int32), ("value", c_int32)]


class GroupbyPlan(Structure):
    _fields_ = [("n_columns", c_uint32), ("column_types", POINTER(c_int32)),
                ("key_columns", POINTER(c_uint32)), ("n_keys", c_uint32),
                ("filter", POINTER(Instr)), ("n_filter", c_uint32),
                ("values", POINTER(Instr)), ("value_lengths", POINTER(c_uint32)), ("n_values", c_uint32),
                ("aggs", POINTER(GroupbyAgg)), ("n_aggs", c_uint32),
                ("strings", POINTER(PString)), ("n_strings", c_uint32)]


class GroupbyOutput(Structure):
    _fields_ = [("first_row", POINTER(c_uint64)), ("key_hash", POINTER(c_uint64)),
                ("row_count", POINTER(c_uint64)), ("agg_f64", POINTER(POINTER(c_double))),
                ("agg_i64", POINTER(POINTER(c_int64))), ("agg_sum128", POINTER(POINTER(c_uint64)))]


class ParquetStats(Structure):
    """prism_parquet_stats."""
    _fields_ = [("present", c_int32), ("physical_type", c_int32), ("null_count", c_int64), ("has_min", c_int32),
                ("has_max", c_int32), ("min", c_void_p), ("min_len", c_uint32), ("max", c_void_p),
                ("max_len", c_uint32)]


class ArrowSchema(Structure):
    pass


ArrowSchema._fields_ = [("format", c_char_p), ("name", c_char_p), ("metadata", c_char_p), ("flags", c_int64),
                        ("n_children", c_int64), ("children", POINTER(POINTER(ArrowSchema))),
                        ("dictionary", POINTER(ArrowSchema)), ("release", c_void_p), ("private_data", c_void_p)]


class ArrowArray(Structure):
    pass


ArrowArray._fields_ = [("length", c_int64), ("null_count", c_int64), ("offset", c_int64), ("n_buffers", c_int64),
                       ("n_children", c_int64), ("buffers", POINTER(c_void_p)),
                       ("children", POINTER(POINTER(ArrowArray))), ("dictionary", POINTER(ArrowArray)),
                       ("release", c_void_p), ("private_data", c_void_p)]


class HashJoinConfig(Structure):
    _fields_ = [("join_type", c_int32), ("build_keys", POINTER(c_uint32)), ("n_build_keys", c_uint32),
                ("probe_keys", POINTER(c_uint32)), ("n_probe_keys", c_uint32)]


class AggExprC(Structure):
    _fields_ = [("column", c_uint32), ("func", c_int32), ("output_name", c_char_p)]


class HashAggConfigC(Structure):
    _fields_ = [("group_by", POINTER(c_uint32)), ("n_group_by", c_uint32), ("aggregates", POINTER(AggExprC)),
                ("n_aggregates", c_uint32), ("flags", c_uint32)]


class Scalar(Structure):
    _fields_ = [("type", c_int32), ("reserved", c_int32), ("i64", c_int64), ("f64", c_double),
                ("str", c_char_p), ("str_len", c_uint32), ("reserved2", c_uint32)]


class PredNode(Structure):
    _fields_ = [("kind", c_int32), ("column", c_uint32), ("left", c_int32), ("right", c_int32),
                ("value", Scalar), ("pattern", c_char_p), ("pattern_len", c_uint32), ("reserved", c_uint32)]


class ExprNode(Structure):
    _fields_ = [("kind", c_int32), ("column", c_uint32), ("value", Scalar), ("op", c_int32), ("left", c_int32),
                ("right", c_int32), ("clause_preds", POINTER(c_int32)), ("clause_exprs", POINTER(c_int32)),
                ("n_clauses", c_uint32), ("else_expr", c_int32)]


class JitStats(Structure):
    _fields_ = [("kernels", c_uint64), ("failures", c_uint64), ("hits", c_uint64), ("compile_ms", c_double)]


class ServiceConfig(Structure):
    _fields_ = [("hbm_budget", c_uint64), ("max_inflight", c_uint32), ("max_batch_requests", c_uint32),
                ("max_batch_rows", c_uint64), ("arena_bytes", c_uint64), ("use_graphs", c_int32),
                ("autotune", c_int32)]


class Request(Structure):
    _fields_ = [("kind", c_int32), ("target", c_uint32), ("columns", POINTER(Column)), ("probe_keys", c_void_p),
                ("probe_payload", c_void_p), ("n_rows", c_uint64), ("out_mask", c_void_p), ("probe_idx", c_void_p),
                ("build_idx", c_void_p), ("pair_capacity", c_uint64)]


class RequestResult(Structure):
    _fields_ = [("status", c_int32), ("kind", c_int32), ("n_rows", c_uint64), ("count", c_uint64),
                ("probe_sum_lo", c_uint64), ("probe_sum_hi", c_int64), ("build_sum_lo", c_uint64),
                ("build_sum_hi", c_int64), ("latency_ms", c_double), ("queue_ms", c_double), ("batch_id", c_uint64)]

    @property
    def probe_sum(self) -> int:
        return (self.probe_sum_hi << 64) + self.probe_sum_lo

    @property
    def build_sum(self) -> int:
        return (self.build_sum_hi << 64) + self.build_sum_lo


class ServiceStats(Structure):
    _fields_ = [("submitted", c_uint64), ("completed", c_uint64), ("batches", c_uint64), ("tables_built", c_uint64),
                ("tables_evicted", c_uint64), ("resident_table_bytes", c_uint64), ("graph_launches", c_uint64),
                ("p50_ms", c_double), ("p99_ms", c_double), ("mean_batch_requests", c_double),
                ("mean_batch_gpu_ms", c_double), ("max_batch_gpu_ms", c_double), ("mean_batch_launch_ms", c_double),
                ("mean_batch_detect_ms", c_double), ("mean_batch_rows", c_double), ("mean_batch_kernel_ms", c_double),
                ("tile_rows", c_uint32 * 3), ("blocks_per_cu", c_uint32 * 3), ("tuned", c_int32 * 3)]


REQ = dict(GROUPBY=1, PROBE=2, FILTER=3)
SCALAR = dict(INT32=0, INT64=1, FLOAT64=2, UTF8=3, BOOLEAN=4, DATE32=5)
PRED = dict(LITERAL=0, EQ=1, NE=2, LT=3, LE=4, GT=5, GE=6, IS_NULL=7, IS_NOT_NULL=8, AND=9, OR=10, NOT=11,
            LIKE=12, ILIKE=13)
EXPR = dict(COLUMN=0, LITERAL=1, BINARY=2, NEGATE=3, IF_THEN=4)
JOIN = dict(INNER=0, LEFT=1, RIGHT=2, FULL=3, LEFT_SEMI=4, LEFT_ANTI=5)
AGG_FLAG_EXACT_DECIMAL = 1

_lib = None

# (name, restype, argtypes) for every function of include/prism_hip.h.
_SIGNATURES = [
    ("prism_last_error", c_char_p, []),
    ("prism_version", c_char_p, []),
    ("prism_ctx_create", c_int32, [c_int, POINTER(c_void_p)]),
    ("prism_ctx_destroy", None, [c_void_p]),
    ("prism_ctx_set_stream", c_int32, [c_void_p, c_void_p]),
    ("prism_ctx_stream", c_void_p, [c_void_p]),
    ("prism_ctx_synchronize", c_int32, [c_void_p]),
    ("prism_ctx_device_bytes", c_uint64, [c_void_p]),
    ("prism_device_alloc", c_int32, [c_void_p, c_uint64, POINTER(c_void_p)]),
    ("prism_device_free", c_int32, [c_void_p, c_void_p]),
    ("prism_copy_to_device", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64]),
    ("prism_copy_to_host", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64]),
    ("prism_copy_device", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64]),
    ("prism_ctx_timer_begin", c_int32, [c_void_p]),
    ("prism_ctx_timer_end", c_int32, [c_void_p, POINTER(c_double)]),
    ("prism_ctx_kernel_timing", c_int32, [c_void_p, c_int32]),
    ("prism_ctx_kernel_timing_read", c_int32, [c_void_p, c_char_p, POINTER(c_double), POINTER(c_uint64)]),
    ("prism_join_build_i64", c_int32, [c_void_p, c_void_p, c_uint64, c_int32, POINTER(c_void_p)]),
    ("prism_join_build_i64_payload", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_int32,
                                               POINTER(c_void_p)]),
    ("prism_join_table_info_get", c_int32, [c_void_p, POINTER(JoinTableInfo)]),
    ("prism_join_build_i64_opts", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, POINTER(BuildOpts),
                                            POINTER(c_void_p)]),
    ("prism_join_tune", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_uint64,
                                  POINTER(BuildOpts), POINTER(TuneCandidate), c_uint32, POINTER(c_uint32)]),
    ("prism_join_table_free", None, [c_void_p]),
    ("prism_join_probe_i64", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_uint64, c_void_p]),
    ("prism_join_probe_i64_noalloc", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_void_p,
                                               c_void_p, c_void_p, c_void_p, c_uint64, c_void_p]),
    ("prism_groupby_create", c_int32, [c_void_p, POINTER(GroupbyPlan), c_uint64, POINTER(c_void_p)]),
    ("prism_groupby_update", c_int32, [c_void_p, POINTER(Column), c_uint64, c_uint64]),
    ("prism_groupby_finish", c_int32, [c_void_p, POINTER(c_uint64)]),
    ("prism_groupby_read", c_int32, [c_void_p, POINTER(GroupbyOutput)]),
    ("prism_groupby_reset", c_int32, [c_void_p]),
    ("prism_groupby_destroy", None, [c_void_p]),
    ("prism_groupby_jit_source", c_int32, [POINTER(GroupbyPlan), c_int32, c_uint32, c_int32, c_char_p, c_uint64,
                                           POINTER(c_uint64)]),
    ("prism_jit_get_stats", c_int32, [POINTER(JitStats)]),
    ("prism_jit_last_error", c_char_p, []),
    ("prism_hash_table_build", c_int32, [c_void_p, c_void_p, c_void_p, POINTER(c_uint32), c_uint32,
                                         POINTER(c_void_p)]),
    ("prism_hash_table_probe", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_uint32), c_uint32,
                                         c_void_p, c_void_p, c_void_p, c_void_p]),
    ("prism_hash_table_free", None, [c_void_p]),
    ("prism_hash_join", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, POINTER(HashJoinConfig),
                                  c_void_p, c_void_p]),
    ("prism_probe_matches_chunked", c_int32, [c_void_p, POINTER(c_void_p), c_uint32, c_void_p, c_void_p, c_void_p,
                                              POINTER(c_uint32), POINTER(c_uint32), c_uint32, c_void_p, c_void_p,
                                              c_void_p]),
    ("prism_hash_join_probe_chunked", c_int32, [c_void_p, POINTER(c_void_p), c_uint32, c_void_p, c_void_p,
                                                c_void_p, POINTER(HashJoinConfig), c_void_p, c_void_p,
                                                POINTER(c_uint32)]),
    ("prism_hash_aggregate", c_int32, [c_void_p, c_void_p, c_void_p, POINTER(HashAggConfigC), c_void_p, c_void_p]),
    ("prism_hash_aggregate_batches", c_int32, [c_void_p, POINTER(c_void_p), c_uint32, c_void_p,
                                               POINTER(HashAggConfigC), c_void_p, c_void_p]),
    ("prism_ipc_read_host", c_int32, [c_void_p, c_uint64, c_void_p, c_uint32, POINTER(c_uint32), c_void_p]),
    ("prism_ipc_read_device", c_int32, [c_void_p, c_void_p, c_uint64, POINTER(c_void_p)]),
    ("prism_ipc_stream_info", c_int32, [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
    ("prism_ipc_field", c_int32, [c_void_p, c_uint32, POINTER(c_char_p), POINTER(c_int32), POINTER(c_int32)]),
    ("prism_ipc_column", c_int32, [c_void_p, c_uint32, c_uint32, POINTER(c_uint64), POINTER(Column)]),
    ("prism_ipc_stream_free", None, [c_void_p]),
    ("prism_ipc_write", c_int32, [c_void_p, c_uint32, c_void_p, POINTER(c_void_p), POINTER(c_uint64)]),
    ("prism_ipc_buffer_free", None, [c_void_p]),
    ("prism_parquet_open", c_int32, [c_void_p, c_uint64, POINTER(c_void_p)]),
    ("prism_parquet_file_info", c_int32, [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_int64)]),
    ("prism_parquet_field", c_int32, [c_void_p, c_uint32, POINTER(c_char_p), POINTER(c_int32), POINTER(c_int32),
                                      POINTER(c_int32)]),
    ("prism_parquet_row_group", c_int32, [c_void_p, c_uint32, POINTER(c_int64)]),
    ("prism_parquet_statistics", c_int32, [c_void_p, c_uint32, c_uint32, POINTER(ParquetStats)]),
    ("prism_parquet_close", None, [c_void_p]),
    ("prism_parquet_read_device", c_int32, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_uint32,
                                            POINTER(c_void_p)]),
    ("prism_parquet_table_info", c_int32, [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
    ("prism_parquet_table_column", c_int32, [c_void_p, c_uint32, c_uint32, POINTER(c_uint64), POINTER(Column)]),
    ("prism_parquet_table_export", c_int32, [c_void_p, c_uint32, c_void_p, c_void_p]),
    ("prism_parquet_table_free", None, [c_void_p]),
    ("prism_expand_worker_aggregates", c_int32, [POINTER(AggExprC), c_uint32, POINTER(AggExprC), POINTER(c_uint32),
                                                 POINTER(c_uint32)]),
    ("prism_merge_partial_aggregates", c_int32, [c_void_p, POINTER(c_void_p), c_uint32, c_void_p, c_uint32,
                                                 POINTER(AggExprC), c_uint32, c_void_p, c_void_p]),
    ("prism_join_agg_create", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, POINTER(c_void_p)]),
    ("prism_join_agg_probe", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64]),
    ("prism_join_agg_finish", c_int32, [c_void_p, POINTER(c_uint64)]),
    ("prism_join_agg_read", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("prism_join_agg_reset", c_int32, [c_void_p]),
    ("prism_join_agg_destroy", None, [c_void_p]),
    ("prism_service_create", c_int32, [c_void_p, POINTER(ServiceConfig), POINTER(c_void_p)]),
    ("prism_service_destroy", None, [c_void_p]),
    ("prism_service_register_table", c_int32, [c_void_p, c_uint32, c_void_p, c_void_p, c_uint64]),
    ("prism_service_put_table", c_int32, [c_void_p, c_char_p, c_void_p, c_void_p, c_uint64, POINTER(c_uint32)]),
    ("prism_service_table_id", c_int32, [c_void_p, c_char_p, POINTER(c_uint32)]),
    ("prism_service_clear_table", c_int32, [c_void_p, c_char_p]),
    ("prism_service_drop_exchange", c_int32, [c_void_p, c_char_p]),
    ("prism_service_register_groupby", c_int32, [c_void_p, c_uint32, POINTER(GroupbyPlan), c_uint64]),
    ("prism_service_register_filter", c_int32, [c_void_p, c_uint32, POINTER(Instr), c_uint32, POINTER(PString),
                                                c_uint32, POINTER(c_int32), c_uint32]),
    ("prism_service_submit", c_int32, [c_void_p, POINTER(Request), POINTER(c_uint64)]),
    ("prism_service_poll", c_int32, [c_void_p, c_uint64, POINTER(c_int32)]),
    ("prism_service_wait", c_int32, [c_void_p, c_uint64, POINTER(RequestResult)]),
    ("prism_service_read_groups", c_int32, [c_void_p, c_uint64, POINTER(GroupbyOutput)]),
    ("prism_service_release", c_int32, [c_void_p, c_uint64]),
    ("prism_service_drain", c_int32, [c_void_p]),
    ("prism_service_stats_get", c_int32, [c_void_p, POINTER(ServiceStats)]),
    ("prism_service_reset_stats", c_int32, [c_void_p]),
    ("prism_evaluate_predicate", c_int32, [c_void_p, c_void_p, c_void_p, POINTER(PredNode), c_uint32, c_void_p,
                                           c_void_p]),
    ("prism_filter_batch", c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("prism_project_batch", c_int32, [c_void_p, c_void_p, c_void_p, POINTER(c_uint32), c_uint32, c_void_p,
                                      c_void_p]),
    ("prism_filter_and_project", c_int32, [c_void_p, c_void_p, c_void_p, POINTER(PredNode), c_uint32,
                                           POINTER(c_uint32), c_uint32, c_void_p, c_void_p]),
    ("prism_evaluate_scalar_expr", c_int32, [c_void_p, c_void_p, c_void_p, POINTER(ExprNode), c_uint32,
                                             POINTER(PredNode), c_uint32, c_int32, c_void_p, c_void_p]),
    ("prism_project_batch_with_exprs", c_int32, [c_void_p, c_void_p, c_void_p, POINTER(c_uint32), c_uint32,
                                                 POINTER(ExprNode), c_uint32, POINTER(c_int32), c_uint32,
                                                 POINTER(PredNode), c_uint32, c_void_p, c_void_p]),
    ("prism_filter_mask", c_int32, [c_void_p, POINTER(Instr), c_uint32, POINTER(PString), c_uint32, POINTER(c_int32),
                                    POINTER(Column), c_uint32, c_uint64, c_void_p, POINTER(c_uint64)]),
    ("prism_string_like", c_int32, [c_void_p, c_void_p, c_void_p, c_char_p, c_uint32, c_int32, c_void_p,
                                    c_void_p]),
    ("prism_datagen_uniform_i64", c_int32, [c_void_p, c_void_p, c_uint64, c_uint64, c_uint32, c_int64,
                                            c_uint64]),
    ("prism_datagen_uniform_i32", c_int32, [c_void_p, c_void_p, c_uint64, c_uint64, c_uint32, c_int64,
                                            c_uint64]),
    ("prism_datagen_bits_i64", c_int32, [c_void_p, c_void_p, c_uint64, c_uint64, c_uint32]),
    ("prism_datagen_permutation_i64", c_int32, [c_void_p, c_void_p, c_uint64, c_uint64, c_uint32]),
    ("prism_datagen_zipf_i64", c_int32, [c_void_p, c_void_p, c_uint64, c_uint64, c_uint32, c_void_p,
                                         c_uint64, c_uint32]),
    ("prism_datagen_lower_utf8", c_int32, [c_void_p, c_void_p, c_uint64, c_void_p, c_uint64, c_uint64, c_uint32,
                                           c_uint32, c_uint32, POINTER(c_uint64)]),
    ("prism_datagen_flag_utf8", c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64, c_uint32,
                                          c_char_p, c_uint32]),
]


def signatures():
    return list(_SIGNATURES)


def load() -> ctypes.CDLL:
    """Load the HIP library once.  Raises RuntimeError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libprism_hip.so is not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != PRISM_OK:
        msg = load().prism_last_error().decode(errors="replace")
        raise PrismError(status, msg)
