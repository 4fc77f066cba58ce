"""Parquet scan (native/prism-executor/src/parquet_scan.rs) over the C ABI:
footer and statistics parsed by prism_parquet_open, row groups skipped by the
reference's statistics rule, projected column chunks decoded into HBM by
prism_parquet_read_device.

`parquet_scan` returns host RecordBatches as the reference does (one per read
row group, exported from HBM); `parquet_scan_device` keeps the decoded columns
resident for the device operators.  URIs are local paths or file:// URIs."""
from __future__ import annotations

import ctypes
import os
import struct
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import pyarrow as pa

from . import _lib as L
from .filter_project import (And, Eq, Ge, Gt, ILike, IsNotNull, IsNull, Le, Like, Literal, Lt, Ne, Not, Or,
                             Predicate, ScalarValue)

_ARROW = {1: pa.int32(), 2: pa.int64(), 3: pa.float64(), 4: pa.utf8(), 5: pa.bool_(), 6: pa.date32()}
_PHYS_BOOLEAN, _PHYS_INT32, _PHYS_INT64, _PHYS_DOUBLE, _PHYS_BYTE_ARRAY = 0, 1, 2, 5, 6


@dataclass
class ParquetScanConfig:
    """ParquetScanConfig (parquet_scan.rs:28-51)."""
    uris: List[str]
    predicate: Optional[Predicate] = None
    projection: Optional[List[int]] = None
    batch_size: int = 8192
    skip_expand: bool = False
    target_schema: Optional[pa.Schema] = None


@dataclass
class ColumnStats:
    """One column chunk's Statistics as the parquet crate exposes them:
    physical type and min / max decoded from their PLAIN bytes (None when
    absent)."""
    physical_type: int
    min: object
    max: object
    null_count: Optional[int]


def _decode_stat(phys: int, raw: bytes):
    if phys == _PHYS_INT32 and len(raw) == 4:
        return struct.unpack("<i", raw)[0]
    if phys == _PHYS_INT64 and len(raw) == 8:
        return struct.unpack("<q", raw)[0]
    if phys == _PHYS_DOUBLE and len(raw) == 8:
        return struct.unpack("<d", raw)[0]
    if phys == _PHYS_BYTE_ARRAY:
        return bytes(raw)
    if phys == _PHYS_BOOLEAN and len(raw) == 1:
        return bool(raw[0])
    return None


class ParquetFile:
    """prism_parquet_file: a parsed footer over `data` (kept referenced here)."""

    def __init__(self, data: bytes):
        self.lib = L.load()  # footer parsing needs no device
        self.data = data
        h = ctypes.c_void_p()
        L.check(self.lib.prism_parquet_open(ctypes.cast(ctypes.c_char_p(data), ctypes.c_void_p), len(data),
                                            ctypes.byref(h)))
        self.handle = h
        ng, nc, nr = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int64()
        L.check(self.lib.prism_parquet_file_info(h, ctypes.byref(ng), ctypes.byref(nc), ctypes.byref(nr)))
        self.num_row_groups, self.num_columns, self.num_rows = ng.value, nc.value, nr.value
        self.fields = []  # (name, PRISM_TYPE_* or 0, nullable, physical type)
        for c in range(self.num_columns):
            name, t, nl, ph = ctypes.c_char_p(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
            L.check(self.lib.prism_parquet_field(h, c, ctypes.byref(name), ctypes.byref(t), ctypes.byref(nl),
                                                 ctypes.byref(ph)))
            self.fields.append((name.value.decode(), t.value, bool(nl.value), ph.value))

    @property
    def schema(self) -> pa.Schema:
        """Arrow schema of the file (leaves the decoder cannot read are typed null)."""
        return pa.schema([pa.field(n, _ARROW.get(t, pa.null()), nl) for n, t, nl, _ in self.fields])

    def row_group_rows(self, rg: int) -> int:
        n = ctypes.c_int64()
        L.check(self.lib.prism_parquet_row_group(self.handle, rg, ctypes.byref(n)))
        return n.value

    def statistics(self, rg: int, col: int) -> Optional[ColumnStats]:
        s = L.ParquetStats()
        L.check(self.lib.prism_parquet_statistics(self.handle, rg, col, ctypes.byref(s)))
        if not s.present:
            return None
        lo = _decode_stat(s.physical_type, ctypes.string_at(s.min, s.min_len)) if s.has_min else None
        hi = _decode_stat(s.physical_type, ctypes.string_at(s.max, s.max_len)) if s.has_max else None
        return ColumnStats(s.physical_type, lo, hi, s.null_count if s.null_count >= 0 else None)

    def close(self) -> None:
        if self.handle:
            self.lib.prism_parquet_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParquetTable:
    """prism_parquet_table: decoded column chunks resident in HBM, one batch per
    row group."""

    def __init__(self, ctx, f: ParquetFile, row_groups: Sequence[int], columns: Sequence[int]):
        self.ctx = ctx
        self.file = f  # the decoded columns do not reference the bytes, the handle does
        rg = (ctypes.c_uint32 * max(len(row_groups), 1))(*row_groups)
        cols = (ctypes.c_uint32 * max(len(columns), 1))(*columns)
        h = ctypes.c_void_p()
        L.check(ctx.lib.prism_parquet_read_device(ctx.handle, f.handle, rg, len(row_groups), cols, len(columns),
                                                  ctypes.byref(h)))
        self.handle = h
        nb, nc = ctypes.c_uint32(), ctypes.c_uint32()
        L.check(ctx.lib.prism_parquet_table_info(h, ctypes.byref(nb), ctypes.byref(nc)))
        self.n_batches, self.n_columns = nb.value, nc.value

    def batch(self, i: int) -> Tuple[int, List[L.Column]]:
        """(rows, device column descriptors) of batch i."""
        cols = []
        n = ctypes.c_uint64()
        for c in range(self.n_columns):
            col = L.Column()
            L.check(self.ctx.lib.prism_parquet_table_column(self.handle, i, c, ctypes.byref(n), ctypes.byref(col)))
            cols.append(col)
        return n.value, cols

    def to_arrow(self, i: int) -> pa.RecordBatch:
        arr, sch = L.ArrowArray(), L.ArrowSchema()
        L.check(self.ctx.lib.prism_parquet_table_export(self.handle, i, ctypes.addressof(arr),
                                                        ctypes.addressof(sch)))
        return pa.RecordBatch._import_from_c(ctypes.addressof(arr), ctypes.addressof(sch))

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.prism_parquet_table_free(self.handle)
            self.handle = None


# ------------------------------------------------------------ file listing
def _local_path(uri: str) -> str:
    if uri.startswith("file://"):
        return uri[len("file://"):]
    if "://" in uri:
        raise NotImplementedError(f"object store scheme of {uri!r} is not supported (local paths only)")
    return uri


def list_files(uri: str, extension: str = "parquet") -> List[str]:
    """object_source.rs:74-103: every object under the prefix ending in
    `.extension`, or the prefix itself when it names one; sorted by path."""
    path = _local_path(uri)
    suffix = "." + extension
    out = []
    if os.path.isdir(path):
        for root, _, files in os.walk(path):
            out.extend(os.path.join(root, f) for f in files if f.endswith(suffix))
    elif os.path.isfile(path) and path.endswith(suffix):
        out.append(path)
    return sorted(out)


# ------------------------------------------------------------ row-group skipping
def _check_range(val, lo, hi, op: str) -> bool:
    """check_range / check_range_str (parquet_scan.rs:471-489)."""
    if op == "eq":
        return lo <= val <= hi
    if op == "lt":
        return lo < val
    if op == "le":
        return lo <= val
    if op == "gt":
        return hi > val
    return hi >= val


_I64_MIN, _I64_MAX, _I32_MIN, _I32_MAX = -(1 << 63), (1 << 63) - 1, -(1 << 31), (1 << 31) - 1
_F64_MAX = 1.7976931348623157e308


def stats_might_match(op: str, val: ScalarValue, stats: ColumnStats) -> bool:
    """stats_might_match (parquet_scan.rs:437-469).  Integer stats compare as
    f64 (`as f64` on both sides), absent bounds default to the type's range."""
    k = val.kind
    if k == "Int64" and stats.physical_type == _PHYS_INT64:
        lo = _I64_MIN if stats.min is None else stats.min
        hi = _I64_MAX if stats.max is None else stats.max
        return _check_range(float(val.value), float(lo), float(hi), op)
    if k == "Int32" and stats.physical_type == _PHYS_INT32:
        lo = _I32_MIN if stats.min is None else stats.min
        hi = _I32_MAX if stats.max is None else stats.max
        return _check_range(float(val.value), float(lo), float(hi), op)
    if k == "Float64" and stats.physical_type == _PHYS_DOUBLE:
        lo = -_F64_MAX if stats.min is None else stats.min
        hi = _F64_MAX if stats.max is None else stats.max
        return _check_range(float(val.value), lo, hi, op)
    if k == "Utf8" and stats.physical_type == _PHYS_BYTE_ARRAY:
        if stats.min is None or stats.max is None:
            return True
        # String::from_utf8_lossy, then str ordering (byte order of the UTF-8)
        lo = stats.min.decode("utf-8", "replace").encode()
        hi = stats.max.decode("utf-8", "replace").encode()
        return _check_range(val.value.encode(), lo, hi, op)
    return True


_OPS = {Eq: "eq", Lt: "lt", Le: "le", Gt: "gt", Ge: "ge"}


def col_stats_might_match(col: int, op: str, val: ScalarValue, f: ParquetFile, rg: int) -> bool:
    """col_stats_might_match (parquet_scan.rs:401-435): the schema field's name
    picks the column chunk; no field, chunk or statistics -> might match."""
    if col >= len(f.fields):
        return True
    name = f.fields[col][0]
    chunk = next((c for c in range(f.num_columns) if f.fields[c][0] == name), None)
    if chunk is None:
        return True
    stats = f.statistics(rg, chunk)
    if stats is None:
        return True
    return stats_might_match(op, val, stats)


def row_group_might_match(predicate: Predicate, f: ParquetFile, rg: int) -> bool:
    """row_group_might_match (parquet_scan.rs:357-390): False only when the
    statistics prove no row of the row group can match."""
    p = predicate
    if isinstance(p, Literal):
        return bool(p.value)
    for cls, op in _OPS.items():
        if type(p) is cls:
            return col_stats_might_match(p.column, op, p.value, f, rg)
    if isinstance(p, (Ne, IsNull, IsNotNull, Not, Like, ILike)):
        return True
    if isinstance(p, And):
        return row_group_might_match(p.left, f, rg) and row_group_might_match(p.right, f, rg)
    if isinstance(p, Or):
        return row_group_might_match(p.left, f, rg) or row_group_might_match(p.right, f, rg)
    raise TypeError(f"not a Predicate: {p!r}")


# ------------------------------------------------------------ schema adaptation
def adapt_batch_to_schema(batch: pa.RecordBatch, target: pa.Schema) -> pa.RecordBatch:
    """adapt_batch_to_schema (parquet_scan.rs:302-355): target order, missing
    columns as nulls, extra columns dropped, differing types cast (a failed
    cast raises)."""
    cols = []
    names = batch.schema.names
    for f in target:
        if f.name in names:
            c = batch.column(names.index(f.name))
            if c.type != f.type:
                try:
                    c = c.cast(f.type)
                except (pa.ArrowInvalid, pa.ArrowNotImplementedError) as e:
                    raise RuntimeError(f"adapt_batch_to_schema: cannot cast column '{f.name}' from {c.type} "
                                       f"to {f.type}: {e}") from e
            cols.append(c)
        else:
            cols.append(pa.nulls(batch.num_rows, f.type))
    return pa.RecordBatch.from_arrays(cols, schema=target)


def expand_projected_batches(batches: List[pa.RecordBatch], proj_cols: Sequence[int],
                             full_schema: pa.Schema) -> List[pa.RecordBatch]:
    """expand_projected_batches (parquet_scan.rs:497-548): back to the full
    schema, non-projected columns as nulls and typed nullable."""
    n_full = len(full_schema)
    proj_map = [None] * n_full
    for bi, oi in enumerate(proj_cols):
        if oi < n_full:
            proj_map[oi] = bi
    fields = [f if proj_map[i] is not None else pa.field(f.name, f.type, True) for i, f in enumerate(full_schema)]
    schema = pa.schema(fields)
    out = []
    for b in batches:
        cols = [b.column(m) if m is not None else pa.nulls(b.num_rows, full_schema.field(i).type)
                for i, m in enumerate(proj_map)]
        out.append(pa.RecordBatch.from_arrays(cols, schema=schema))
    return out


# ------------------------------------------------------------ scan
def _resolve(uris: Sequence[str]) -> List[str]:
    files = []
    for u in uris:
        got = list_files(u)
        if not got:
            raise RuntimeError(f"no .parquet files found at {u}")
        files.extend(got)
    return files


def _read(path: str) -> bytes:
    with open(path, "rb") as fh:
        return fh.read()


def _plan(f: ParquetFile, config: ParquetScanConfig) -> Tuple[List[int], List[int], int]:
    rgs = [g for g in range(f.num_row_groups)
           if config.predicate is None or row_group_might_match(config.predicate, f, g)]
    cols = list(range(f.num_columns))
    if config.projection is not None:
        valid = [c for c in config.projection if c < f.num_columns]
        if valid:
            cols = sorted(set(valid))  # ProjectionMask::roots keeps file order
    return rgs, cols, f.num_row_groups - len(rgs)


def parquet_scan_device(config: ParquetScanConfig, ctx) -> List[ParquetTable]:
    """The scan with the decoded columns left in HBM (one table per file)."""
    out = []
    for path in _resolve(config.uris) if config.uris else []:
        f = ParquetFile(_read(path))
        rgs, cols, _ = _plan(f, config)
        if rgs:
            out.append(ParquetTable(ctx, f, rgs, cols))
    return out


def parquet_scan(config: ParquetScanConfig, ctx) -> List[pa.RecordBatch]:
    """parquet_scan (parquet_scan.rs:55-159) returning one RecordBatch per read
    row group, files in path order."""
    if not config.uris:
        return []
    files = _resolve(config.uris)
    if config.target_schema is not None:
        full_schema = config.target_schema
    else:
        full_schema = ParquetFile(_read(files[0])).schema
    batches = []
    for path in files:
        f = ParquetFile(_read(path))
        rgs, cols, _ = _plan(f, config)
        if not rgs:
            continue
        t = ParquetTable(ctx, f, rgs, cols)
        try:
            for i in range(t.n_batches):
                b = t.to_arrow(i)
                if b.num_rows == 0:
                    continue
                if config.target_schema is not None and config.projection is None:
                    b = adapt_batch_to_schema(b, config.target_schema)
                batches.append(b)
        finally:
            t.close()
    if not config.skip_expand and config.projection is not None and len(config.projection) < len(full_schema):
        return expand_projected_batches(batches, config.projection, full_schema)
    return batches


def parquet_row_count(uris: Sequence[str], predicate: Optional[Predicate] = None) -> int:
    """parquet_row_count (parquet_scan.rs:260-286): metadata only."""
    total = 0
    for u in uris:
        for path in list_files(u):
            f = ParquetFile(_read(path))
            for g in range(f.num_row_groups):
                if predicate is None or row_group_might_match(predicate, f, g):
                    total += f.row_group_rows(g)
    return total
