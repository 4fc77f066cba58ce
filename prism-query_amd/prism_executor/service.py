"""Operator service: device-resident tables and plans, continuous batching of
group-by / probe / filter requests (prism_service_* in libprism_hip.so).

The plain counterpart of each request kind is the one-request-per-launch
operator (device.GroupBy, device.JoinTable, prism_filter_mask); results are
identical.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_int32, c_uint64, c_void_p
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .device import Context, build_plan, instrs, ptr, read_groups


class Service:
    def __init__(self, ctx: Context, hbm_budget: int = 0, max_inflight: int = 3, max_batch_requests: int = 256,
                 max_batch_rows: int = 0, arena_bytes: int = 0, use_graphs: int = 1, autotune: bool = True):
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = L.ServiceConfig(hbm_budget, max_inflight, max_batch_requests, max_batch_rows, arena_bytes,
                              int(use_graphs), 1 if autotune else 0)
        self.handle = c_void_p()
        L.check(self.lib.prism_service_create(ctx.handle, byref(cfg), byref(self.handle)))
        self._n_aggs = {}
        self._keep = []

    def register_table(self, table_id: int, keys: np.ndarray, payload: Optional[np.ndarray] = None) -> None:
        k = np.ascontiguousarray(keys, np.int64)
        p = None if payload is None else np.ascontiguousarray(payload, np.int32)
        L.check(self.lib.prism_service_register_table(self.handle, table_id, k.ctypes.data,
                                                      None if p is None else p.ctypes.data, len(k)))

    def put_table(self, key: str, keys: np.ndarray, payload: Optional[np.ndarray] = None) -> int:
        """Append rows to the resident table under a store key (PartitionStore::put,
        prism-flight shuffle_writer.rs:52-55); returns the table id for probes."""
        k = np.ascontiguousarray(keys, np.int64)
        p = None if payload is None else np.ascontiguousarray(payload, np.int32)
        tid = ctypes.c_uint32()
        L.check(self.lib.prism_service_put_table(self.handle, key.encode(), k.ctypes.data,
                                                 None if p is None else p.ctypes.data, len(k), byref(tid)))
        return tid.value

    def table_id(self, key: str) -> int:
        tid = ctypes.c_uint32()
        L.check(self.lib.prism_service_table_id(self.handle, key.encode(), byref(tid)))
        return tid.value

    def clear_table(self, key: str) -> None:
        """PartitionStore::clear (shuffle_writer.rs:62-65)."""
        L.check(self.lib.prism_service_clear_table(self.handle, key.encode()))

    def drop_exchange(self, exchange_id: str) -> None:
        """PartitionStore::drop_exchange (shuffle_writer.rs:77-85): every table
        under "exchange/<exchange_id>/"."""
        L.check(self.lib.prism_service_drop_exchange(self.handle, exchange_id.encode()))

    def register_groupby(self, plan_id: int, column_types, keys, values, aggs, filter_prog=None, strings=(),
                         max_groups: int = 1024) -> None:
        plan, keep = build_plan(column_types, keys, values, aggs, filter_prog, strings)
        L.check(self.lib.prism_service_register_groupby(self.handle, plan_id, byref(plan), max_groups))
        self._n_aggs[plan_id] = len(aggs)

    def register_filter(self, plan_id: int, prog, column_types, strings=()) -> None:
        ins = instrs(prog)
        code = (L.Instr * len(ins))(*ins)
        strs = [bytes(x) for x in strings]
        st = (L.PString * max(len(strs), 1))(*[L.PString(x, len(x)) for x in strs])
        ct = (ctypes.c_int32 * max(len(column_types), 1))(*column_types)
        L.check(self.lib.prism_service_register_filter(self.handle, plan_id, code, len(ins), st, len(strs), ct,
                                                       len(column_types)))

    def _submit(self, req: L.Request) -> int:
        t = c_uint64()
        L.check(self.lib.prism_service_submit(self.handle, byref(req), byref(t)))
        return t.value

    def submit_groupby(self, plan_id: int, columns: Sequence[L.Column], n_rows: int) -> int:
        cols = (L.Column * max(len(columns), 1))(*columns)
        return self._submit(L.Request(L.REQ["GROUPBY"], plan_id, cols, None, None, n_rows, None))

    def submit_probe(self, table_id: int, keys, payload=None, n_rows: Optional[int] = None, probe_idx=None,
                     build_idx=None, capacity: Optional[int] = None) -> int:
        """HashTable::probe against a resident table: counts and exact payload
        sums, and with probe_idx / build_idx (device u32 arrays) the matched
        pairs in probe order, at most `capacity` of them (default: n_rows)."""
        n = keys.numel() if n_rows is None else n_rows
        cap = n if capacity is None else capacity
        return self._submit(L.Request(L.REQ["PROBE"], table_id, None, ptr(keys), ptr(payload), n, None,
                                      ptr(probe_idx), ptr(build_idx), cap if probe_idx is not None else 0))

    def submit_filter(self, plan_id: int, columns: Sequence[L.Column], n_rows: int, out_mask=None) -> int:
        cols = (L.Column * max(len(columns), 1))(*columns)
        return self._submit(L.Request(L.REQ["FILTER"], plan_id, cols, None, None, n_rows, ptr(out_mask)))

    def poll(self, ticket: int) -> bool:
        d = c_int32()
        L.check(self.lib.prism_service_poll(self.handle, ticket, byref(d)))
        return bool(d.value)

    def wait(self, ticket: int) -> L.RequestResult:
        r = L.RequestResult()
        L.check(self.lib.prism_service_wait(self.handle, ticket, byref(r)))
        return r

    def read_groups(self, ticket: int, plan_id: int, n_groups: int) -> dict:
        return read_groups(lambda out: L.check(self.lib.prism_service_read_groups(self.handle, ticket, out)),
                           n_groups, self._n_aggs[plan_id])

    def release(self, ticket: int) -> None:
        L.check(self.lib.prism_service_release(self.handle, ticket))

    def drain(self) -> None:
        L.check(self.lib.prism_service_drain(self.handle))

    def stats(self) -> dict:
        s = L.ServiceStats()
        L.check(self.lib.prism_service_stats_get(self.handle, byref(s)))
        return {"submitted": s.submitted, "completed": s.completed, "batches": s.batches,
                "tables_built": s.tables_built, "tables_evicted": s.tables_evicted,
                "resident_table_bytes": s.resident_table_bytes, "graph_launches": s.graph_launches,
                "p50_ms": s.p50_ms, "p99_ms": s.p99_ms, "mean_batch_requests": s.mean_batch_requests,
                "mean_batch_gpu_ms": s.mean_batch_gpu_ms, "max_batch_gpu_ms": s.max_batch_gpu_ms,
                "mean_batch_launch_ms": s.mean_batch_launch_ms, "mean_batch_detect_ms": s.mean_batch_detect_ms,
                "mean_batch_rows": s.mean_batch_rows, "mean_batch_kernel_ms": s.mean_batch_kernel_ms,
                "tile_rows": list(s.tile_rows), "blocks_per_cu": list(s.blocks_per_cu), "tuned": list(s.tuned)}

    def reset_stats(self) -> None:
        L.check(self.lib.prism_service_reset_stats(self.handle))

    def close(self) -> None:
        if self.handle:
            self.lib.prism_service_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def filter_mask(ctx: Context, prog, column_types, columns: Sequence[L.Column], n_rows: int, out_mask=None,
                strings=()) -> int:
    """One-request-per-launch filter (prism_filter_mask): selected row count."""
    ins = instrs(prog)
    code = (L.Instr * len(ins))(*ins)
    strs = [bytes(x) for x in strings]
    st = (L.PString * max(len(strs), 1))(*[L.PString(x, len(x)) for x in strs])
    ct = (ctypes.c_int32 * max(len(column_types), 1))(*column_types)
    cols = (L.Column * max(len(columns), 1))(*columns)
    sel = c_uint64()
    L.check(ctx.lib.prism_filter_mask(ctx.handle, code, len(ins), st, len(strs), ct, cols, len(columns), n_rows,
                                      ptr(out_mask), byref(sel)))
    return sel.value
