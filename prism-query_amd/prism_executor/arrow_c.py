"""pyarrow <-> Arrow C Data Interface structs for the C ABI's Arrow layer."""
from __future__ import annotations

import ctypes

import pyarrow as pa

from . import _lib as L


class Exported:
    """A RecordBatch exported to C structs (the producer side keeps it alive)."""

    def __init__(self, batch: pa.RecordBatch):
        self.batch = batch
        self.array = L.ArrowArray()
        self.schema = L.ArrowSchema()
        batch._export_to_c(ctypes.addressof(self.array), ctypes.addressof(self.schema))

    @property
    def aptr(self) -> int:
        return ctypes.addressof(self.array)

    @property
    def sptr(self) -> int:
        return ctypes.addressof(self.schema)

    def __del__(self):
        for s in (self.array, self.schema):
            if s.release:
                ctypes.CFUNCTYPE(None, ctypes.c_void_p)(s.release)(ctypes.addressof(s))


def export_schema(schema: pa.Schema) -> L.ArrowSchema:
    s = L.ArrowSchema()
    schema._export_to_c(ctypes.addressof(s))
    return s


def release_schema(s: L.ArrowSchema) -> None:
    if s.release:
        ctypes.CFUNCTYPE(None, ctypes.c_void_p)(s.release)(ctypes.addressof(s))


def import_batch(array: L.ArrowArray, schema: L.ArrowSchema) -> pa.RecordBatch:
    """Take ownership of a C struct array + schema produced by the library."""
    return pa.RecordBatch._import_from_c(ctypes.addressof(array), ctypes.addressof(schema))


def import_array(array: L.ArrowArray, schema: L.ArrowSchema) -> pa.Array:
    return pa.Array._import_from_c(ctypes.addressof(array), ctypes.addressof(schema))


def import_array_with_schema(array: L.ArrowArray, arrow_type: pa.DataType) -> pa.Array:
    return pa.Array._import_from_c(ctypes.addressof(array), arrow_type)


def u32s(values) -> ctypes.Array:
    values = list(values)
    return (ctypes.c_uint32 * max(len(values), 1))(*values)
