"""Arrow IPC streams at the JNI boundary through the C ABI: ingest (the byte
arrays prism-jni decodes with arrow's StreamReader, native/prism-jni/src/lib.rs:96-121)
as Arrow C Data arrays over the caller's bytes (no copy) or straight into HBM,
and the reply stream prism-jni writes with arrow's StreamWriter (lib.rs:111-121)."""
from __future__ import annotations

import ctypes
from typing import List, Sequence, Tuple

import pyarrow as pa

from . import _lib as L
from .arrow_c import import_batch


def _buf(data: bytes):
    return ctypes.cast(ctypes.c_char_p(data), ctypes.c_void_p)


def read_ipc_host(data: bytes) -> Tuple[List[pa.RecordBatch], bytes]:
    """The stream's record batches; their buffers point into `data`, which is
    returned beside them and must stay referenced while they are used."""
    lib = L.load()  # no device needed
    n = ctypes.c_uint32()
    cap = 1
    while True:
        arrays = (L.ArrowArray * cap)()
        schema = L.ArrowSchema()
        st = lib.prism_ipc_read_host(_buf(data), len(data), arrays, cap, ctypes.byref(n), ctypes.addressof(schema))
        if st == 3 and b"batches, the output room" in lib.prism_last_error():
            cap *= 4
            continue
        L.check(st)
        break
    sch = pa.Schema._import_from_c(ctypes.addressof(schema))  # takes the C schema over
    out = [pa.RecordBatch._import_from_c(ctypes.addressof(arrays[i]), sch) for i in range(n.value)]
    return out, data


class IpcStream:
    """prism_ipc_stream: a stream decoded into HBM (device columns per batch)."""

    def __init__(self, ctx, data: bytes):
        self.ctx = ctx
        h = ctypes.c_void_p()
        L.check(ctx.lib.prism_ipc_read_device(ctx.handle, _buf(data), len(data), ctypes.byref(h)))
        self.handle = h
        nb, nc = ctypes.c_uint32(), ctypes.c_uint32()
        L.check(ctx.lib.prism_ipc_stream_info(h, ctypes.byref(nb), ctypes.byref(nc)))
        self.n_batches, self.n_columns = nb.value, nc.value
        self.fields = []
        for c in range(self.n_columns):
            name, t, nl = ctypes.c_char_p(), ctypes.c_int32(), ctypes.c_int32()
            L.check(ctx.lib.prism_ipc_field(h, c, ctypes.byref(name), ctypes.byref(t), ctypes.byref(nl)))
            self.fields.append((name.value.decode(), t.value, bool(nl.value)))

    def batch(self, i: int) -> Tuple[int, List[L.Column]]:
        """(rows, device column descriptors) of record batch i."""
        cols = []
        n = ctypes.c_uint64()
        for c in range(self.n_columns):
            col = L.Column()
            L.check(self.ctx.lib.prism_ipc_column(self.handle, i, c, ctypes.byref(n), ctypes.byref(col)))
            cols.append(col)
        return n.value, cols

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.prism_ipc_stream_free(self.handle)
            self.handle = None


def write_ipc(batches: Sequence[pa.RecordBatch]) -> bytes:
    """The Arrow IPC stream of the batches (one schema, the batches, the
    end-of-stream marker); empty when there are no batches, as the reference's
    reply is (prism-jni/src/lib.rs:111-121)."""
    lib = L.load()  # no device needed
    n = len(batches)
    arrays = (L.ArrowArray * max(n, 1))()
    schema = L.ArrowSchema()
    exported = []
    try:
        if n:
            batches[0].schema._export_to_c(ctypes.addressof(schema))
            exported.append(schema)
            for i, b in enumerate(batches):
                if not b.schema.equals(batches[0].schema):
                    raise ValueError("record batches of one stream share one schema")
                b._export_to_c(ctypes.addressof(arrays[i]))
                exported.append(arrays[i])
        out, size = ctypes.c_void_p(), ctypes.c_uint64()
        L.check(lib.prism_ipc_write(arrays, n, ctypes.addressof(schema), ctypes.byref(out), ctypes.byref(size)))
        try:
            return ctypes.string_at(out, size.value) if size.value else b""
        finally:
            lib.prism_ipc_buffer_free(out)
    finally:
        for s in exported:
            if s.release:
                ctypes.CFUNCTYPE(None, ctypes.c_void_p)(s.release)(ctypes.addressof(s))
