"""Oracle (test infrastructure only): a CPU restatement of the reference's
filter / project operators, native/prism-executor/src/filter_project.rs:20-404
and string_ops.rs:16-25, over pyarrow RecordBatches.

Only tests, the smoke check and the benchmark's CPU baseline use this module;
no product code imports it.  Trees are the dataclasses of
prism_executor.filter_project (plain data, no device code).

Semantics restated from the reference and from arrow 57.3.0, which the
reference calls for the element kernels:
  * compare_column (filter_project.rs:346-390): the accepted column/scalar
    type pairs; anything else is InvalidArgument.
  * arrow-ord cmp on floats: IEEE total order (f64::total_cmp).
  * arrow boolean `and` / `or` / `not`: null when any input is null.
  * LIKE: '%' any run, '_' one character, '\\' escapes (string_ops.rs:16-19
    -> arrow_string::like); ILIKE folds ASCII letters.
  * filter (FilterBuilder): null mask entries drop the row.
  * BinaryOp / Negate: operands cast to Float64 (filter_project.rs:236-254).
  * IfThen: result typed as the first THEN branch; clauses applied last to
    first with zip, so the first true clause wins (filter_project.rs:256-292).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pyarrow as pa

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "prism-query_amd"))
from prism_executor import filter_project as fp  # noqa: E402  (tree dataclasses only)

import oracle as orc  # noqa: E402


class OracleError(Exception):
    def __init__(self, kind: str, msg: str):
        super().__init__(msg)
        self.kind = kind


_TYPE_DEBUG = {pa.int32(): "Int32", pa.int64(): "Int64", pa.float64(): "Float64", pa.utf8(): "Utf8",
               pa.bool_(): "Boolean", pa.date32(): "Date32"}


def total_key(x: np.ndarray) -> np.ndarray:
    """f64::total_cmp as an int64 key: negative values get every bit but the
    sign flipped (core::f64::total_cmp)."""
    b = np.ascontiguousarray(x, dtype=np.float64).view(np.int64)
    return b ^ ((b >> 63).astype(np.uint64) >> np.uint64(1)).astype(np.int64)


def _valid(arr: pa.Array) -> np.ndarray:
    return np.asarray(arr.is_valid(), dtype=bool)


def _ints(arr: pa.Array) -> np.ndarray:
    if pa.types.is_date32(arr.type):
        arr = arr.cast(pa.int32())
    return np.asarray(arr.fill_null(0), dtype=np.int64)


def _cmp(op: str, a, b) -> np.ndarray:
    return {"Eq": a == b, "Ne": a != b, "Lt": a < b, "Le": a <= b, "Gt": a > b, "Ge": a >= b}[op]


def _strings(arr: pa.Array):
    return [None if v is None else v.encode() for v in arr.to_pylist()]


def eval_pred(batch: pa.RecordBatch, p) -> tuple:
    """(values, valid) boolean numpy arrays (evaluate_predicate, filter_project.rs:65-123)."""
    n = batch.num_rows
    if isinstance(p, fp.Literal):
        return np.full(n, bool(p.value)), np.ones(n, bool)
    if isinstance(p, fp._Cmp):
        col = batch.column(p.column)
        t, k = col.type, p.value.kind
        op = type(p).__name__
        ok = ((t == pa.int32() and k == "Int32") or (t == pa.int64() and k == "Int64")
              or (t == pa.date32() and k in ("Date32", "Int32")))
        if ok:
            return _cmp(op, _ints(col), np.int64(p.value.value)), _valid(col)
        if t == pa.float64() and k == "Float64":
            a = total_key(np.asarray(col.fill_null(0.0), np.float64))
            b = total_key(np.array([p.value.value], np.float64))[0]
            return _cmp(op, a, b), _valid(col)
        if t == pa.utf8() and k == "Utf8":
            lit = p.value.value.encode()
            vals = np.array([False if s is None else bool(_cmp(op, s, lit)) for s in _strings(col)], bool)
            return vals, _valid(col)
        raise OracleError("InvalidArgument", f"Cannot compare column type {_TYPE_DEBUG.get(t, t)} with scalar "
                                             f"{k}({p.value.value})")
    if isinstance(p, (fp.IsNull, fp.IsNotNull)):
        v = _valid(batch.column(p.column))
        return (~v if isinstance(p, fp.IsNull) else v), np.ones(n, bool)
    if isinstance(p, (fp.And, fp.Or)):
        lv, lm = eval_pred(batch, p.left)
        rv, rm = eval_pred(batch, p.right)
        v = (lv & rv) if isinstance(p, fp.And) else (lv | rv)
        return v, lm & rm
    if isinstance(p, fp.Not):
        v, m = eval_pred(batch, p.inner)
        return ~v, m
    if isinstance(p, (fp.Like, fp.ILike)):
        col = batch.column(p.column)
        if col.type != pa.utf8():
            what = "LIKE" if isinstance(p, fp.Like) else "ILIKE"
            raise OracleError("InvalidArgument", f"{what} requires Utf8 column, got {_TYPE_DEBUG.get(col.type)}")
        strs = _strings(col)
        res = orc.like([s if s is not None else b"" for s in strs], p.pattern.encode(), isinstance(p, fp.ILike))
        return np.asarray(res, bool) & _valid(col), _valid(col)
    raise TypeError(p)


def mask_array(values: np.ndarray, valid: np.ndarray) -> pa.BooleanArray:
    return pa.array([bool(v) if m else None for v, m in zip(values, valid)], pa.bool_())


def filter_rows(batch: pa.RecordBatch, values: np.ndarray, valid: np.ndarray) -> pa.RecordBatch:
    """filter_batch (filter_project.rs:126-135): null = drop."""
    idx = np.nonzero(values & valid)[0]
    return batch.take(pa.array(idx, pa.int64()))


def _as_f64(arr: pa.Array) -> tuple:
    t = arr.type
    if t in (pa.int32(), pa.int64(), pa.bool_()):
        return np.asarray(arr.cast(pa.float64()).fill_null(0.0), np.float64), _valid(arr)
    if t == pa.float64():
        return np.asarray(arr.fill_null(0.0), np.float64), _valid(arr)
    raise OracleError("Arrow", f"Cast error: Casting from {_TYPE_DEBUG.get(t, t)} to Float64 not supported")


def _literal(v, n: int) -> pa.Array:
    typ = {"Int32": pa.int32(), "Int64": pa.int64(), "Float64": pa.float64(), "Utf8": pa.utf8(),
           "Boolean": pa.bool_(), "Date32": pa.date32()}[v.kind]
    return pa.array([v.value] * n, typ)


def _cast_to(arr: pa.Array, t: pa.DataType) -> pa.Array:
    if arr.type == t:
        return arr
    if t == pa.float64() and arr.type in (pa.int32(), pa.int64(), pa.bool_()):
        return arr.cast(t)
    if t == pa.int64() and arr.type in (pa.int32(), pa.bool_(), pa.date32()):
        return arr.cast(pa.int32()).cast(t) if arr.type == pa.date32() else arr.cast(t)
    if t == pa.int32() and arr.type in (pa.bool_(), pa.date32()):
        return arr.cast(t)
    if t == pa.date32() and arr.type == pa.int32():
        return arr.cast(t)
    if t == pa.bool_() and arr.type in (pa.int32(), pa.int64()):
        return arr.cast(t)
    raise OracleError("Arrow", f"Cast error: Casting from {arr.type} to {t} not supported")


def eval_expr(batch: pa.RecordBatch, e) -> pa.Array:
    """evaluate_scalar_expr (filter_project.rs:216-294)."""
    n = batch.num_rows
    if isinstance(e, fp.ColumnRef):
        return batch.column(e.index)
    if isinstance(e, fp.LiteralExpr):
        return _literal(e.value, n)
    if isinstance(e, fp.BinaryOp):
        a, am = _as_f64(eval_expr(batch, e.left))
        b, bm = _as_f64(eval_expr(batch, e.right))
        with np.errstate(all="ignore"):
            r = {0: a + b, 1: a - b, 2: a * b, 3: a / b}[int(e.op)]
        m = am & bm
        return pa.array(r, pa.float64(), mask=~m)
    if isinstance(e, fp.Negate):
        a, am = _as_f64(eval_expr(batch, e.inner))
        return pa.array(-a, pa.float64(), mask=~am)
    if isinstance(e, fp.IfThen):
        if not e.clauses:
            if e.else_expr is None:
                raise OracleError("InvalidArgument", "IfThen expression requires at least one clause or an else branch")
            return eval_expr(batch, e.else_expr)
        branches = [eval_expr(batch, x) for _, x in e.clauses]
        target = branches[0].type
        result = eval_expr(batch, e.else_expr) if e.else_expr is not None else pa.nulls(n, target)
        result = _cast_to(result, target)
        for (p, _), br in reversed(list(zip(e.clauses, branches))):
            v, m = eval_pred(batch, p)
            take = v & m
            br = _cast_to(br, target)
            result = pa.array([br[i].as_py() if take[i] else result[i].as_py() for i in range(n)], target)
        return result
    raise TypeError(e)


def project_with_exprs(batch: pa.RecordBatch, indices, exprs) -> pa.RecordBatch:
    """project_batch_with_exprs (filter_project.rs:305-333)."""
    arrays = [batch.column(i) for i in indices]
    fields = [batch.schema.field(i) for i in indices]
    for k, x in enumerate(exprs):
        a = eval_expr(batch, x)
        arrays.append(a)
        fields.append(pa.field(f"expr_{k}", a.type, nullable=True))
    return pa.RecordBatch.from_arrays(arrays, schema=pa.schema(fields))
