"""Device layer of the C ABI over HBM buffers allocated through the library.

No other GPU runtime is loaded in the process: torch ships its own HIP/HSA
runtime, and two runtimes in one process do not share a device.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_uint64, c_void_p
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


def ptr(t) -> Optional[int]:
    return None if t is None else t.ptr


class Context:
    """prism_ctx: a device, a stream, scratch memory."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        self.handle = c_void_p()
        L.check(self.lib.prism_ctx_create(device, byref(self.handle)))
        self.device = device

    def set_stream(self, stream_ptr: Optional[int]) -> None:
        L.check(self.lib.prism_ctx_set_stream(self.handle, stream_ptr))

    @property
    def stream(self) -> int:
        return self.lib.prism_ctx_stream(self.handle) or 0

    def sync(self) -> None:
        L.check(self.lib.prism_ctx_synchronize(self.handle))

    def device_bytes(self) -> int:
        return int(self.lib.prism_ctx_device_bytes(self.handle))

    def timer_begin(self) -> None:
        L.check(self.lib.prism_ctx_timer_begin(self.handle))

    def timer_end(self) -> float:
        ms = ctypes.c_double()
        L.check(self.lib.prism_ctx_timer_end(self.handle, byref(ms)))
        return ms.value

    def kernel_timing(self, enable: bool) -> None:
        """Bracket the library's main kernels with HIP events (see the
        prism_ctx_kernel_timing names) until disabled."""
        L.check(self.lib.prism_ctx_kernel_timing(self.handle, 1 if enable else 0))

    def kernel_time(self, name: str) -> tuple:
        """(summed ms, launches) of one timed kernel since the last read."""
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        L.check(self.lib.prism_ctx_kernel_timing_read(self.handle, name.encode(), byref(ms), byref(n)))
        return ms.value, n.value

    def close(self) -> None:
        if self.handle:
            self.lib.prism_ctx_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A 1-D typed buffer in HBM (allocated by prism_device_alloc)."""

    def __init__(self, ctx: Context, n: int, dtype, _base: "DeviceArray" = None, _start: int = 0):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.size = int(n)
        self._base = _base
        if _base is None:
            p = c_void_p()
            L.check(ctx.lib.prism_device_alloc(ctx.handle, max(self.size * self.dtype.itemsize, 1), byref(p)))
            self._alloc = p.value
            self.ptr = p.value
        else:
            self._alloc = None
            self.ptr = _base.ptr + _start * self.dtype.itemsize

    @classmethod
    def from_numpy(cls, ctx: Context, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        if a.size:
            L.check(ctx.lib.prism_copy_to_device(ctx.handle, d.ptr, a.ctypes.data, a.nbytes))
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.size, self.dtype)
        if self.size:
            L.check(self.ctx.lib.prism_copy_to_host(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def numel(self) -> int:
        return self.size

    def __len__(self) -> int:
        return self.size

    def __getitem__(self, sl) -> "DeviceArray":
        start, stop, step = sl.indices(self.size)
        assert step == 1, "only contiguous views"
        return DeviceArray(self.ctx, max(stop - start, 0), self.dtype, _base=self, _start=start)

    def free(self) -> None:
        if self._alloc:
            self.ctx.lib.prism_device_free(self.ctx.handle, self._alloc)
            self._alloc = None

    def __del__(self):  # pragma: no cover
        try:
            self.free()
        except Exception:
            pass


def empty(ctx: Context, n: int, dtype) -> DeviceArray:
    return DeviceArray(ctx, n, dtype)


def to_device(ctx: Context, a) -> DeviceArray:
    return DeviceArray.from_numpy(ctx, a)


# ---------------------------------------------------------------------------
# Column descriptors
# ---------------------------------------------------------------------------
def column(type_code: int, data, offsets=None, validity=None, offset: int = 0) -> L.Column:
    c = L.Column()
    c.type = type_code
    c.offset = offset
    c.data = ptr(data)
    c.offsets = ptr(offsets)
    c.validity = ptr(validity)
    return c


def col_i64(t, validity=None):
    return column(L.TYPE_INT64, t, validity=validity)


def col_i32(t, validity=None):
    return column(L.TYPE_INT32, t, validity=validity)


def col_f64(t, validity=None):
    return column(L.TYPE_FLOAT64, t, validity=validity)


def col_date32(t, validity=None):
    return column(L.TYPE_DATE32, t, validity=validity)


def col_utf8(data, offsets, validity=None):
    return column(L.TYPE_UTF8, data, offsets=offsets, validity=validity)


# ---------------------------------------------------------------------------
# Join
# ---------------------------------------------------------------------------
class JoinTable:
    """HashTable::build on a device-resident int64 key column (hash_join.rs:87-141)."""

    def __init__(self, ctx: Context, keys, kind: int = L.TABLE_AUTO, payload=None, opts: Optional[L.BuildOpts] = None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.handle = c_void_p()
        if opts is not None:  # layout and probe form chosen by the caller (or by tune())
            L.check(self.lib.prism_join_build_i64_opts(ctx.handle, ptr(keys), ptr(payload), keys.numel(), byref(opts),
                                                       byref(self.handle)))
        elif payload is None:
            L.check(self.lib.prism_join_build_i64(ctx.handle, ptr(keys), keys.numel(), kind, byref(self.handle)))
        else:
            L.check(self.lib.prism_join_build_i64_payload(ctx.handle, ptr(keys), ptr(payload), keys.numel(), kind,
                                                          byref(self.handle)))
        self.info = L.JoinTableInfo()
        L.check(self.lib.prism_join_table_info_get(self.handle, byref(self.info)))

    def probe_async(self, keys, result_buf, probe_payload=None, build_payload=None, out_probe=None,
                    out_build=None, capacity: int = 0, probe_row_base: int = 0) -> None:
        L.check(self.lib.prism_join_probe_i64(self.ctx.handle, self.handle, ptr(keys), keys.numel(),
                                              probe_row_base, ptr(probe_payload), ptr(build_payload),
                                              ptr(out_probe), ptr(out_build), capacity, ptr(result_buf)))

    def probe(self, keys, probe_payload=None, build_payload=None, want_pairs: bool = True,
              capacity: Optional[int] = None, probe_row_base: int = 0):
        """HashTable::probe (hash_join.rs:143-200).  Returns (probe_idx, build_idx, ProbeResult);
        the index tensors are None when want_pairs is False."""
        ctx = self.ctx
        res = DeviceArray(ctx, ctypes.sizeof(L.ProbeResult), np.uint8)
        n = keys.numel()
        cap = (n if self.info.unique else max(n, 1024)) if capacity is None else capacity
        while True:
            op = ob = None
            if want_pairs:
                op = DeviceArray(ctx, max(cap, 1), np.uint32)
                ob = DeviceArray(ctx, max(cap, 1), np.uint32)
            self.probe_async(keys, res, probe_payload, build_payload, op, ob, cap if want_pairs else 0,
                             probe_row_base)
            self.ctx.sync()
            r = L.ProbeResult.from_buffer_copy(res.numpy().tobytes())
            if r.error:
                raise L.PrismError(4, "Execution error: probe look-back wait gave up")
            if want_pairs and r.matches > cap:
                cap = int(r.matches)
                continue
            if want_pairs:
                return op[: int(r.matches)], ob[: int(r.matches)], r
            return None, None, r

    @staticmethod
    def tune(ctx: Context, keys, payload=None, sample=None, sample_payload=None):
        """prism_join_tune: (fastest BuildOpts, [TuneCandidate, ...]) for these
        build keys, timed on one probe of `sample`."""
        best = L.BuildOpts()
        cands = (L.TuneCandidate * 32)()
        n = ctypes.c_uint32()
        L.check(ctx.lib.prism_join_tune(ctx.handle, ptr(keys), ptr(payload), keys.numel(), ptr(sample),
                                        ptr(sample_payload), sample.numel() if sample is not None else 0,
                                        byref(best), cands, 32, byref(n)))
        return best, [cands[i] for i in range(n.value)]

    def close(self) -> None:
        if self.handle:
            self.lib.prism_join_table_free(self.handle)
            self.handle = c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class JoinAggregate:
    """Join -> GROUP BY build attribute, COUNT(*), SUM(probe payload) without
    materialising pairs (the fused join-aggregate path, prism-substrait
    executor.rs:333-458).  Build arrays must stay alive until finish()."""

    def __init__(self, ctx: Context, build_keys, build_attr):
        self.ctx = ctx
        self.lib = ctx.lib
        self.handle = c_void_p()
        self._keep = (build_keys, build_attr)
        L.check(self.lib.prism_join_agg_create(ctx.handle, ptr(build_keys), ptr(build_attr), build_keys.numel(),
                                               byref(self.handle)))

    def probe(self, keys, payload=None, probe_row_base: int = 0) -> None:
        L.check(self.lib.prism_join_agg_probe(self.handle, ptr(keys), ptr(payload), keys.numel(), probe_row_base))

    def finish(self) -> int:
        n = c_uint64()
        L.check(self.lib.prism_join_agg_finish(self.handle, byref(n)))
        return n.value

    def read(self, n_groups: int) -> dict:
        k = max(n_groups, 1)
        attr = np.zeros(k, np.int32)
        cnt = np.zeros(k, np.uint64)
        lo = np.zeros(k, np.uint64)
        hi = np.zeros(k, np.int64)
        first = np.zeros(k, np.uint64)
        L.check(self.lib.prism_join_agg_read(self.handle, attr.ctypes.data, cnt.ctypes.data, lo.ctypes.data,
                                             hi.ctypes.data, first.ctypes.data))
        sums = [int(h) * (1 << 64) + int(l) for l, h in zip(lo[:n_groups].tolist(), hi[:n_groups].tolist())]
        return {"attr": attr[:n_groups], "count": cnt[:n_groups], "sum": sums,
                "first_probe_row": (first[:n_groups] >> np.uint64(32)).astype(np.int64),
                "first_build_row": (first[:n_groups] & np.uint64(0xFFFFFFFF)).astype(np.int64)}

    def reset(self) -> None:
        L.check(self.lib.prism_join_agg_reset(self.handle))

    def close(self) -> None:
        if self.handle:
            self.lib.prism_join_agg_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Group-by
# ---------------------------------------------------------------------------
def instrs(prog: Sequence[tuple]) -> list:
    """[(op_name, arg, i64, f64), ...] -> list of Instr."""
    out = []
    for t in prog:
        op = t[0]
        ins = L.Instr()
        ins.op = L.OP[op] if isinstance(op, str) else op
        ins.arg = t[1] if len(t) > 1 else 0
        ins.i64 = t[2] if len(t) > 2 else 0
        ins.f64 = t[3] if len(t) > 3 else 0.0
        out.append(ins)
    return out


def build_plan(column_types: Sequence[int], keys: Sequence[int], values: Sequence[Sequence[tuple]],
               aggs: Sequence[tuple], filter_prog=None, strings: Sequence[bytes] = ()):
    """prism_groupby_plan + the ctypes arrays it points to (keep them alive)."""
    ct = (ctypes.c_int32 * max(len(column_types), 1))(*column_types)
    kc = (ctypes.c_uint32 * max(len(keys), 1))(*keys)
    flat, lens = [], []
    for v in values:
        ins = instrs(v)
        flat += ins
        lens.append(len(ins))
    vals = (L.Instr * max(len(flat), 1))(*flat)
    vl = (ctypes.c_uint32 * max(len(lens), 1))(*lens)
    ag = (L.GroupbyAgg * max(len(aggs), 1))(*[L.GroupbyAgg(L.AGG[f] if isinstance(f, str) else f, v)
                                              for f, v in aggs])
    fl = instrs(filter_prog) if filter_prog else []
    fa = (L.Instr * max(len(fl), 1))(*fl)
    strs = [bytes(x) for x in strings]
    st = (L.PString * max(len(strs), 1))(*[L.PString(x, len(x)) for x in strs])
    plan = L.GroupbyPlan(len(column_types), ct, kc, len(keys), fa if fl else None, len(fl), vals, vl,
                         len(lens), ag, len(aggs), st, len(strs))
    return plan, (ct, kc, vals, vl, ag, fa, st, strs)


def jit_source(column_types: Sequence[int], keys: Sequence[int], values: Sequence[Sequence[tuple]],
               aggs: Sequence[tuple], filter_prog=None, strings: Sequence[bytes] = (), form: int = 0,
               nullable: int = 0, compile: bool = False) -> str:
    """Source of the generated group-by update kernel of this plan (form 0 =
    wave pre-aggregation, 1 = per-row slots; bit c of `nullable` = column c has
    a validity bitmap).  compile=True also compiles it for gfx950 (no device
    needed) and raises PrismError with the compiler log when that fails."""
    lib = L.load()
    plan, keep = build_plan(column_types, keys, values, aggs, filter_prog, strings)
    n = c_uint64(0)
    L.check(lib.prism_groupby_jit_source(byref(plan), form, nullable, 0, None, 0, byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    L.check(lib.prism_groupby_jit_source(byref(plan), form, nullable, 1 if compile else 0, buf, n.value + 1,
                                         byref(n)))
    del keep
    return buf.value.decode()


def jit_stats() -> dict:
    """Counters of the kernels generated at run time in this process."""
    lib = L.load()
    s = L.JitStats()
    L.check(lib.prism_jit_get_stats(byref(s)))
    return {"kernels": s.kernels, "failures": s.failures, "hits": s.hits, "compile_ms": s.compile_ms,
            "last_error": (lib.prism_jit_last_error() or b"").decode(errors="replace")}


def read_groups(read_fn, n_groups: int, n_aggs: int) -> dict:
    """Call read_fn(prism_groupby_output*) and return the groups as numpy arrays
    (int128 sums as Python ints)."""
    n = max(n_groups, 1)
    first = np.zeros(n, np.uint64)
    khash = np.zeros(n, np.uint64)
    rcount = np.zeros(n, np.uint64)
    f64 = [np.zeros(n, np.float64) for _ in range(n_aggs)]
    i64 = [np.zeros(n, np.int64) for _ in range(n_aggs)]
    s128 = [np.zeros(2 * n, np.uint64) for _ in range(n_aggs)]
    P = ctypes.POINTER
    fp = (P(ctypes.c_double) * max(n_aggs, 1))(*[a.ctypes.data_as(P(ctypes.c_double)) for a in f64])
    ip = (P(ctypes.c_int64) * max(n_aggs, 1))(*[a.ctypes.data_as(P(ctypes.c_int64)) for a in i64])
    sp = (P(ctypes.c_uint64) * max(n_aggs, 1))(*[a.ctypes.data_as(P(ctypes.c_uint64)) for a in s128])
    out = L.GroupbyOutput(first.ctypes.data_as(P(c_uint64)), khash.ctypes.data_as(P(c_uint64)),
                          rcount.ctypes.data_as(P(c_uint64)), fp, ip, sp)
    read_fn(byref(out))
    sums = []
    for a in s128:
        lo = a[0::2][:n_groups].astype(object)
        hi = a[1::2][:n_groups].view(np.int64).astype(object)
        sums.append([int(h) * (1 << 64) + int(l) for l, h in zip(lo, hi)])
    return {"first_row": first[:n_groups], "key_hash": khash[:n_groups], "row_count": rcount[:n_groups],
            "f64": [a[:n_groups] for a in f64], "i64": [a[:n_groups] for a in i64], "sum128": sums}


class GroupBy:
    """prism_groupby: resident group state fed batch by batch."""

    def __init__(self, ctx: Context, column_types: Sequence[int], keys: Sequence[int],
                 values: Sequence[Sequence[tuple]], aggs: Sequence[tuple], filter_prog=None,
                 strings: Sequence[bytes] = (), expected_groups: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        self.n_aggs = len(aggs)
        plan, self._keep = build_plan(column_types, keys, values, aggs, filter_prog, strings)
        self.handle = c_void_p()
        L.check(self.lib.prism_groupby_create(ctx.handle, byref(plan), expected_groups, byref(self.handle)))

    def update(self, columns: Sequence[L.Column], n_rows: int, row_base: int = 0) -> None:
        arr = (L.Column * max(len(columns), 1))(*columns)
        L.check(self.lib.prism_groupby_update(self.handle, arr, n_rows, row_base))

    def reset(self) -> None:
        L.check(self.lib.prism_groupby_reset(self.handle))

    def finish(self) -> int:
        n = c_uint64()
        L.check(self.lib.prism_groupby_finish(self.handle, byref(n)))
        return int(n.value)

    def read(self, n_groups: int) -> dict:
        return read_groups(lambda out: L.check(self.lib.prism_groupby_read(self.handle, out)), n_groups, self.n_aggs)

    def close(self) -> None:
        if self.handle:
            self.lib.prism_groupby_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
