"""hash_aggregate: the API of native/prism-executor/src/hash_aggregate.rs on the GPU."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Sequence

import pyarrow as pa

from . import _lib as L
from .arrow_c import Exported, export_schema, import_batch, release_schema, u32s
from .context import context


class AggFunc(IntEnum):
    """AggFunc (hash_aggregate.rs:79-86)."""
    Sum = 0
    Count = 1
    Avg = 2
    Min = 3
    Max = 4
    CountDistinct = 5


@dataclass
class AggExpr:
    """AggExpr (hash_aggregate.rs:88-94)."""
    column: int
    func: AggFunc
    output_name: str


@dataclass
class HashAggConfig:
    """HashAggConfig (hash_aggregate.rs:96-101)."""
    group_by: Sequence[int]
    aggregates: Sequence[AggExpr] = field(default_factory=list)


def _cfg(config: HashAggConfig, exact: bool):
    names = [a.output_name.encode() for a in config.aggregates]
    aggs = (L.AggExprC * max(len(names), 1))(*[L.AggExprC(a.column, int(a.func), n)
                                                for a, n in zip(config.aggregates, names)])
    gb = u32s(config.group_by)
    c = L.HashAggConfigC(gb, len(config.group_by), aggs, len(config.aggregates),
                         L.AGG_FLAG_EXACT_DECIMAL if exact else 0)
    return c, (names, aggs, gb)


def hash_aggregate(batch: pa.RecordBatch, config: HashAggConfig, exact: bool = False) -> pa.RecordBatch:
    """hash_aggregate (hash_aggregate.rs:174-176).  exact=True returns SUM of
    integer columns as Decimal128(38, 0) (not in the reference)."""
    return hash_aggregate_batches([batch], config, exact=exact, schema=batch.schema)


def hash_aggregate_batches(batches: Sequence[pa.RecordBatch], config: HashAggConfig, exact: bool = False,
                           schema: pa.Schema = None) -> pa.RecordBatch:
    """hash_aggregate_batches (hash_aggregate.rs:182-339)."""
    ctx = context()
    ex = [Exported(b) for b in batches]
    if schema is None:
        schema = batches[0].schema if batches else pa.schema([])
    sch = export_schema(schema)
    cfg, keep = _cfg(config, exact)
    ptrs = (ctypes.c_void_p * max(len(ex), 1))(*[e.aptr for e in ex])
    out_a, out_s = L.ArrowArray(), L.ArrowSchema()
    try:
        L.check(ctx.lib.prism_hash_aggregate_batches(ctx.handle, ptrs, len(ex), ctypes.addressof(sch),
                                                     ctypes.byref(cfg), ctypes.addressof(out_a),
                                                     ctypes.addressof(out_s)))
    finally:
        release_schema(sch)
    return import_batch(out_a, out_s)
