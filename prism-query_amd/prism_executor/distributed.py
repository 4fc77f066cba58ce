"""Distributed aggregate of prism-bench (native/prism-bench/src/handler.rs:270-344,
397-526) over the GPU operators: producers run the plan with the expanded
aggregate list and ship their partial batches (Flight stays on the host); the
reducer merges them with one device aggregation and rebuilds the original
aggregates."""
from __future__ import annotations

import ctypes
from typing import List, Sequence, Tuple

import pyarrow as pa

from . import _lib as L
from .arrow_c import Exported, export_schema, import_batch, release_schema
from .context import context
from .hash_aggregate import AggExpr, AggFunc, HashAggConfig


def _c_aggs(aggs: Sequence[AggExpr]):
    names = [a.output_name.encode() for a in aggs]
    arr = (L.AggExprC * max(len(aggs), 1))(*[L.AggExprC(a.column, int(a.func), n) for a, n in zip(aggs, names)])
    return arr, names


def expand_worker_aggregates(original: Sequence[AggExpr]) -> Tuple[List[AggExpr], List[List[int]]]:
    """expand_worker_aggregates (handler.rs:397-428): AVG becomes SUM + COUNT of
    the same column, every worker aggregate is named agg_<index>; returns the
    worker list and, per original aggregate, its worker indices."""
    lib = L.load()  # no device needed
    arr, keep = _c_aggs(original)
    out = (L.AggExprC * max(2 * len(original), 1))()
    n = ctypes.c_uint32()
    m = (ctypes.c_uint32 * max(2 * len(original), 1))()
    L.check(lib.prism_expand_worker_aggregates(arr, len(original), out, ctypes.byref(n), m))
    worker = [AggExpr(out[i].column, AggFunc(out[i].func), out[i].output_name.decode()) for i in range(n.value)]
    mapping = [[m[2 * i]] + ([m[2 * i + 1]] if m[2 * i + 1] != 0xFFFFFFFF else []) for i in range(len(original))]
    return worker, mapping


def build_final_merge_config(group_by_count: int, worker_aggs: Sequence[AggExpr]) -> HashAggConfig:
    """build_final_merge_config (handler.rs:430-457): partial SUM / COUNT /
    COUNT DISTINCT columns merge by SUM, MIN by MIN, MAX by MAX."""
    aggs = []
    for i, a in enumerate(worker_aggs):
        if a.func == AggFunc.Avg:
            raise ValueError("worker aggregate list must not contain AVG after expansion")
        f = a.func if a.func in (AggFunc.Min, AggFunc.Max) else AggFunc.Sum
        aggs.append(AggExpr(group_by_count + i, f, a.output_name))
    return HashAggConfig(list(range(group_by_count)), aggs)


def merge_partial_aggregates(partials: Sequence[pa.RecordBatch], group_by_count: int,
                             original: Sequence[AggExpr], schema: pa.Schema = None) -> pa.RecordBatch:
    """Reducer side of execute_distributed_aggregate (handler.rs:311-343): the
    partial batches (group columns, then the worker aggregates of `original`)
    merged on the device, then reconstruct_original_aggregate_batch
    (handler.rs:459-526)."""
    ctx = context()
    if schema is None:
        schema = partials[0].schema
    ex = [Exported(b) for b in partials]
    sch = export_schema(schema)
    arr, keep = _c_aggs(original)
    ptrs = (ctypes.c_void_p * max(len(ex), 1))(*[e.aptr for e in ex])
    out_a, out_s = L.ArrowArray(), L.ArrowSchema()
    try:
        L.check(ctx.lib.prism_merge_partial_aggregates(ctx.handle, ptrs, len(ex), ctypes.addressof(sch),
                                                       group_by_count, arr, len(original),
                                                       ctypes.addressof(out_a), ctypes.addressof(out_s)))
    finally:
        release_schema(sch)
    return import_batch(out_a, out_s)
