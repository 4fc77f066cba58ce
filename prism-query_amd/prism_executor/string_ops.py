"""string_ops: SQL LIKE / ILIKE of native/prism-executor/src/string_ops.rs:16-25
on the GPU."""
from __future__ import annotations

import ctypes

import pyarrow as pa

from . import _lib as L
from .arrow_c import import_array
from .context import context


def _like(array: pa.StringArray, pattern: str, ci: bool) -> pa.BooleanArray:
    ctx = context()
    aa, as_ = L.ArrowArray(), L.ArrowSchema()
    array._export_to_c(ctypes.addressof(aa), ctypes.addressof(as_))
    try:
        b = pattern.encode()
        oa, os_ = L.ArrowArray(), L.ArrowSchema()
        L.check(ctx.lib.prism_string_like(ctx.handle, ctypes.addressof(aa), ctypes.addressof(as_), b, len(b),
                                          1 if ci else 0, ctypes.addressof(oa), ctypes.addressof(os_)))
        return import_array(oa, os_)
    finally:
        for s in (aa, as_):
            if s.release:
                ctypes.CFUNCTYPE(None, ctypes.c_void_p)(s.release)(ctypes.addressof(s))


def string_like(array: pa.StringArray, pattern: str) -> pa.BooleanArray:
    """string_like (string_ops.rs:16-19)."""
    return _like(array, pattern, False)


def string_ilike(array: pa.StringArray, pattern: str) -> pa.BooleanArray:
    """string_ilike (string_ops.rs:21-25): ASCII case folding."""
    return _like(array, pattern, True)
