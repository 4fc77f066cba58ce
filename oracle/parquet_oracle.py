"""Oracle (test infrastructure only): a CPU restatement of the reference's
row-group skipping, native/prism-executor/src/parquet_scan.rs:357-494, over
the statistics that pyarrow's own Parquet reader reports (an independent
reading of the footer).  Only tests use this module.

The reference consults the `parquet` crate's Statistics:
  * Int64 / Int32 / Double statistics against Int64 / Int32 / Float64 scalars
    compare as f64 (`as f64` on the value and both bounds), absent bounds
    default to the type's minimum / maximum (parquet_scan.rs:439-456);
  * ByteArray statistics against a Utf8 scalar compare as strings after
    from_utf8_lossy; an absent bound means "might match" (457-466);
  * every other scalar / statistics pairing, Ne, IsNull, IsNotNull, Not, Like
    and ILike might match; And / Or combine; Literal is itself (357-390);
  * the schema field's name picks the column chunk; no field, no chunk or no
    statistics might match (401-435).
"""
from __future__ import annotations

from typing import Callable, Optional, Sequence, Tuple

_I64 = (-(1 << 63), (1 << 63) - 1)
_I32 = (-(1 << 31), (1 << 31) - 1)
_F64 = (-1.7976931348623157e308, 1.7976931348623157e308)

# (physical type name, min, max) as pyarrow reports it, or None
Stats = Optional[Tuple[str, object, object]]


def check_range(val, lo, hi, op: str) -> bool:
    """parquet_scan.rs:471-489 (the same table serves numbers and strings)."""
    return {"Eq": lambda: lo <= val <= hi, "Lt": lambda: lo < val, "Le": lambda: lo <= val,
            "Gt": lambda: hi > val, "Ge": lambda: hi >= val}[op]()


def stats_might_match(op: str, kind: str, value, st: Stats) -> bool:
    """parquet_scan.rs:437-469."""
    phys, lo, hi = st
    if kind == "Int64" and phys == "INT64":
        return check_range(float(value), float(_I64[0] if lo is None else lo), float(_I64[1] if hi is None else hi),
                           op)
    if kind == "Int32" and phys == "INT32":
        return check_range(float(value), float(_I32[0] if lo is None else lo), float(_I32[1] if hi is None else hi),
                           op)
    if kind == "Float64" and phys == "DOUBLE":
        return check_range(float(value), _F64[0] if lo is None else lo, _F64[1] if hi is None else hi, op)
    if kind == "Utf8" and phys == "BYTE_ARRAY":
        if lo is None or hi is None:
            return True
        as_text = lambda b: b if isinstance(b, str) else bytes(b).decode("utf-8", "replace")
        # Rust str ordering is the byte order of the UTF-8 encoding
        return check_range(value.encode(), as_text(lo).encode(), as_text(hi).encode(), op)
    return True


def row_group_might_match(p, field_names: Sequence[str], stats_of: Callable[[str], Stats]) -> bool:
    """parquet_scan.rs:357-390 + 401-435.  `stats_of(name)` gives the named
    column chunk's statistics (None: no chunk or no statistics)."""
    k = type(p).__name__
    if k == "Literal":
        return bool(p.value)
    if k in ("Eq", "Lt", "Le", "Gt", "Ge"):
        if p.column >= len(field_names):
            return True
        st = stats_of(field_names[p.column])
        if st is None:
            return True
        return stats_might_match(k, p.value.kind, p.value.value, st)
    if k in ("Ne", "IsNull", "IsNotNull", "Not", "Like", "ILike"):
        return True
    if k == "And":
        return row_group_might_match(p.left, field_names, stats_of) and row_group_might_match(p.right, field_names,
                                                                                          stats_of)
    if k == "Or":
        return row_group_might_match(p.left, field_names, stats_of) or row_group_might_match(p.right, field_names,
                                                                                         stats_of)
    raise TypeError(k)


def _physical(v):
    """pyarrow reports DATE statistics as dates; the crate sees the INT32 days."""
    if hasattr(v, "toordinal"):
        return v.toordinal() - 719163  # date(1970, 1, 1).toordinal()
    return v


def pyarrow_stats(md, rg: int) -> Callable[[str], Stats]:
    """stats_of over pyarrow's FileMetaData for row group rg."""
    g = md.row_group(rg)

    def stats_of(name: str) -> Stats:
        for j in range(g.num_columns):
            c = g.column(j)
            if c.path_in_schema == name:
                s = c.statistics
                if s is None:
                    return None
                if not s.has_min_max:
                    return (s.physical_type, None, None)
                return (s.physical_type, _physical(s.min), _physical(s.max))
        return None

    return stats_of
