"""ctypes wrapper of libref_cpu.so: the reference's CPU operator path restated
with its own data structures and parallel decomposition (64-partition
build, probe parallel over batches, chunked aggregation; see ref_cpu.h).
Test infrastructure: timed by bench.py's CPU-baseline legs and checked
against the oracle by tests/test_ref_cpu.py.

timed() follows the reference's scaling bench (benches/scaling.rs:77-93):
2 warm-up runs, then the median of the timed runs.
"""
from __future__ import annotations

import ctypes
import os
import platform
import statistics
import time
from ctypes import POINTER, Structure, c_double, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libref_cpu.so")
BATCH_ROWS = 1_048_576  # the worker's Parquet batch (prism-bench handler.rs:923)


class Groups(Structure):
    _fields_ = [("n_groups", c_int64), ("key_hash", c_uint64 * 1024), ("first_row", c_int64 * 1024),
                ("agg", (c_double * 1024) * 4)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"reference-path CPU library not built at {LIB_PATH}")
        lib = ctypes.CDLL(LIB_PATH)
        lib.rc_join_build.restype = c_void_p
        lib.rc_join_build.argtypes = [c_void_p, c_int64, c_int32]
        lib.rc_join_free.restype = None
        lib.rc_join_free.argtypes = [c_void_p]
        lib.rc_probe_matches_chunked.restype = c_int64
        lib.rc_probe_matches_chunked.argtypes = [c_void_p, c_void_p, c_int64, c_int64, c_int32, POINTER(c_uint64),
                                                 POINTER(c_uint64)]
        lib.rc_groupby_i64.restype = c_int32
        lib.rc_groupby_i64.argtypes = [c_void_p, c_void_p, c_int64, c_int64, c_int32, POINTER(Groups)]
        lib.rc_filter_project_aggregate.restype = c_int32
        lib.rc_filter_project_aggregate.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                    c_void_p, c_int64, c_int32, c_int32, c_int64, c_int32,
                                                    POINTER(Groups)]
        lib.rc_join_aggregate.restype = c_int64
        lib.rc_join_aggregate.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int32, c_int32,
                                          c_void_p, c_void_p]
        _lib = lib
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


class JoinTable:
    """HashTable::build (hash_join.rs:87-141) on an Int64 key column."""

    def __init__(self, keys, threads):
        self.keys = _c(keys, np.int64)
        self.h = load().rc_join_build(self.keys.ctypes.data, len(self.keys), threads)

    def probe_chunked(self, keys, threads, batch_rows=BATCH_ROWS):
        """probe_matches_chunked: (pairs, probe-index checksum, build-index checksum)."""
        k = _c(keys, np.int64)
        pc, bc = c_uint64(), c_uint64()
        n = load().rc_probe_matches_chunked(self.h, k.ctypes.data, len(k), batch_rows, threads, ctypes.byref(pc),
                                            ctypes.byref(bc))
        return int(n), pc.value, bc.value

    def join_aggregate(self, build_attr, probe_key, probe_payload, n_attr, threads, batch_rows=BATCH_ROWS):
        """The executor's fused join -> aggregate: (pairs, count[attr], sum[attr]) as f64."""
        a = _c(build_attr, np.int32)
        k, p = _c(probe_key, np.int64), _c(probe_payload, np.int64)
        cnt = np.zeros(n_attr, np.float64)
        sm = np.zeros(n_attr, np.float64)
        pairs = load().rc_join_aggregate(self.h, a.ctypes.data, k.ctypes.data, p.ctypes.data, len(k), batch_rows,
                                         threads, n_attr, cnt.ctypes.data, sm.ctypes.data)
        return int(pairs), cnt, sm

    def close(self):
        if self.h:
            load().rc_join_free(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        self.close()


def _groups(g: Groups, n_aggs: int) -> dict:
    n = min(int(g.n_groups), 1024)
    return {"n_groups": int(g.n_groups), "key_hash": list(g.key_hash)[:n], "first_row": list(g.first_row)[:n],
            "agg": [list(g.agg[a])[:n] for a in range(n_aggs)]}


def groupby_i64(key, value, threads, batch_rows=BATCH_ROWS) -> dict:
    """hash_aggregate_batches: SUM, COUNT, MIN, MAX of value by key."""
    k, v = _c(key, np.int64), _c(value, np.int64)
    g = Groups()
    load().rc_groupby_i64(k.ctypes.data, v.ctypes.data, len(k), batch_rows, threads, ctypes.byref(g))
    return _groups(g, 4)


def filter_project_aggregate(date, price, discount, flag_a, flag_b, date_max, threads,
                             batch_rows=BATCH_ROWS, variant=0) -> dict:
    """Filter(date <= date_max) -> Project -> Aggregate(SUM, COUNT, AVG of
    price * (100 - discount) (variant 0) or price * 0.01 * (1 - discount * 0.01)
    (variant 1) by (flag_a, flag_b)); flags are (bytes, offsets)."""
    d, p, di = _c(date, np.int32), _c(price, np.int64), _c(discount, np.int64)
    ab, ao = _c(flag_a[0], np.uint8), _c(flag_a[1], np.int32)
    bb, bo = _c(flag_b[0], np.uint8), _c(flag_b[1], np.int32)
    g = Groups()
    load().rc_filter_project_aggregate(d.ctypes.data, p.ctypes.data, di.ctypes.data, ab.ctypes.data, ao.ctypes.data,
                                       bb.ctypes.data, bo.ctypes.data, len(d), date_max, variant, batch_rows,
                                       threads, ctypes.byref(g))
    return _groups(g, 3)


def host_threads() -> int:
    """Host threads to use: every core the process may run on, at most 16
    (a GPU box gives one GPU's job a 16-core share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:  # pragma: no cover
        pass
    return platform.processor() or "unknown"


def timed(fn, runs: int = 5, warmup: int = 2) -> tuple:
    """(median seconds, all timed runs) after `warmup` untimed runs."""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), ts
