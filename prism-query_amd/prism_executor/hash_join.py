"""hash_join: the API of native/prism-executor/src/hash_join.rs on the GPU.

Same names and argument meaning as the reference: probe = left, build = right;
ProbeMatches holds UInt32 index arrays; errors raise PrismError with the
reference's variant names.  Every call runs in libprism_hip.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from enum import IntEnum
from typing import List, Sequence, Tuple

import pyarrow as pa

from . import _lib as L
from .arrow_c import Exported, import_array, import_array_with_schema, import_batch, release_schema, u32s
from .context import context


class JoinType(IntEnum):
    """JoinType (hash_join.rs:48-56)."""
    Inner = 0
    Left = 1
    Right = 2
    Full = 3
    LeftSemi = 4
    LeftAnti = 5


@dataclass
class HashJoinConfig:
    """HashJoinConfig (hash_join.rs:58-64)."""
    join_type: JoinType
    build_keys: Sequence[int]
    probe_keys: Sequence[int]

    def c(self):
        bk, pk = u32s(self.build_keys), u32s(self.probe_keys)
        cfg = L.HashJoinConfig(int(self.join_type), bk, len(self.build_keys), pk, len(self.probe_keys))
        return cfg, (bk, pk)


@dataclass
class ProbeMatches:
    """ProbeMatches (hash_join.rs:66-70)."""
    probe_indices: pa.UInt32Array
    build_indices: pa.UInt32Array


class HashTable:
    """HashTable (hash_join.rs:77-201): the build side resident in HBM."""

    def __init__(self, handle: ctypes.c_void_p, ctx):
        self._h = handle
        self._ctx = ctx

    @staticmethod
    def build(batch: pa.RecordBatch, key_cols: Sequence[int]) -> "HashTable":
        ctx = context()
        e = Exported(batch)
        h = ctypes.c_void_p()
        keys = u32s(key_cols)
        L.check(ctx.lib.prism_hash_table_build(ctx.handle, e.aptr, e.sptr, keys, len(key_cols), ctypes.byref(h)))
        return HashTable(h, ctx)

    def probe(self, probe_batch: pa.RecordBatch, probe_key_cols: Sequence[int]) -> Tuple[pa.UInt32Array, pa.UInt32Array]:
        e = Exported(probe_batch)
        pa_, ps, ba, bs = L.ArrowArray(), L.ArrowSchema(), L.ArrowArray(), L.ArrowSchema()
        keys = u32s(probe_key_cols)
        L.check(self._ctx.lib.prism_hash_table_probe(self._ctx.handle, self._h, e.aptr, e.sptr, keys,
                                                     len(probe_key_cols), ctypes.addressof(pa_), ctypes.addressof(ps),
                                                     ctypes.addressof(ba), ctypes.addressof(bs)))
        return import_array(pa_, ps), import_array(ba, bs)

    def __del__(self):
        if getattr(self, "_h", None):
            self._ctx.lib.prism_hash_table_free(self._h)
            self._h = None


def hash_join(probe_batch: pa.RecordBatch, build_batch: pa.RecordBatch, config: HashJoinConfig) -> pa.RecordBatch:
    """hash_join (hash_join.rs:312-365)."""
    ctx = context()
    p, b = Exported(probe_batch), Exported(build_batch)
    cfg, keep = config.c()
    out_a, out_s = L.ArrowArray(), L.ArrowSchema()
    L.check(ctx.lib.prism_hash_join(ctx.handle, p.aptr, p.sptr, b.aptr, b.sptr, ctypes.byref(cfg),
                                    ctypes.addressof(out_a), ctypes.addressof(out_s)))
    return import_batch(out_a, out_s)


def _batch_ptrs(exported: List[Exported]):
    return (ctypes.c_void_p * max(len(exported), 1))(*[e.aptr for e in exported])


def probe_matches_chunked(probe_batches: Sequence[pa.RecordBatch], build_batch: pa.RecordBatch,
                          probe_key_cols: Sequence[int], build_key_cols: Sequence[int]) -> List[ProbeMatches]:
    """probe_matches_chunked (hash_join.rs:286-309)."""
    ctx = context()
    ex = [Exported(b) for b in probe_batches]
    if not ex:
        return []
    b = Exported(build_batch)
    n = len(ex)
    pa_out = (L.ArrowArray * n)()
    ba_out = (L.ArrowArray * n)()
    sch = L.ArrowSchema()
    L.check(ctx.lib.prism_probe_matches_chunked(ctx.handle, _batch_ptrs(ex), n, ex[0].sptr, b.aptr, b.sptr,
                                                u32s(probe_key_cols), u32s(build_key_cols), len(probe_key_cols),
                                                ctypes.addressof(pa_out), ctypes.addressof(ba_out),
                                                ctypes.addressof(sch)))
    release_schema(sch)  # one UInt32 schema; pyarrow takes the type per import
    return [ProbeMatches(import_array_with_schema(pa_out[i], pa.uint32()),
                         import_array_with_schema(ba_out[i], pa.uint32())) for i in range(n)]


def hash_join_probe_chunked(probe_batches: Sequence[pa.RecordBatch], build_batch: pa.RecordBatch,
                            config: HashJoinConfig) -> List[pa.RecordBatch]:
    """hash_join_probe_chunked (hash_join.rs:206-282)."""
    ctx = context()
    ex = [Exported(b) for b in probe_batches]
    if not ex:
        return []
    b = Exported(build_batch)
    cfg, keep = config.c()
    n = len(ex)
    outs = (L.ArrowArray * n)()
    sch = L.ArrowSchema()
    n_out = ctypes.c_uint32()
    L.check(ctx.lib.prism_hash_join_probe_chunked(ctx.handle, _batch_ptrs(ex), n, ex[0].sptr, b.aptr, b.sptr,
                                                  ctypes.byref(cfg), ctypes.addressof(outs), ctypes.addressof(sch),
                                                  ctypes.byref(n_out)))
    if not n_out.value:
        return []
    schema = pa.Schema._import_from_c(ctypes.addressof(sch))  # takes ownership of `sch`
    return [pa.RecordBatch._import_from_c(ctypes.addressof(outs[i]), schema) for i in range(n_out.value)]
