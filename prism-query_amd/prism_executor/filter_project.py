"""filter_project: the API of native/prism-executor/src/filter_project.rs on the GPU.

Predicate / ScalarValue / ScalarExpr mirror the reference enums
(filter_project.rs:21-62,185-213); the functions keep the reference's names,
argument meaning, output schemas ("expr_{i}" columns) and error kinds.  Trees
are flattened into node arrays and evaluated by libprism_hip.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Optional, Sequence, Tuple, Union

import pyarrow as pa

from . import _lib as L
from .arrow_c import Exported, import_array, import_batch, u32s
from .context import context


# ---------------------------------------------------------------- ScalarValue
@dataclass(frozen=True)
class ScalarValue:
    """ScalarValue (filter_project.rs:54-62): kind is one of Int32, Int64,
    Float64, Utf8, Boolean, Date32."""
    kind: str
    value: Union[int, float, str, bool]

    @staticmethod
    def Int32(v: int) -> "ScalarValue":
        return ScalarValue("Int32", int(v))

    @staticmethod
    def Int64(v: int) -> "ScalarValue":
        return ScalarValue("Int64", int(v))

    @staticmethod
    def Float64(v: float) -> "ScalarValue":
        return ScalarValue("Float64", float(v))

    @staticmethod
    def Utf8(v: str) -> "ScalarValue":
        return ScalarValue("Utf8", str(v))

    @staticmethod
    def Boolean(v: bool) -> "ScalarValue":
        return ScalarValue("Boolean", bool(v))

    @staticmethod
    def Date32(v: int) -> "ScalarValue":
        return ScalarValue("Date32", int(v))


_SCALAR_CODE = {"Int32": 0, "Int64": 1, "Float64": 2, "Utf8": 3, "Boolean": 4, "Date32": 5}


# ------------------------------------------------------------------ Predicate
class Predicate:
    """Predicate (filter_project.rs:21-51)."""


@dataclass
class Literal(Predicate):
    value: bool


@dataclass
class _Cmp(Predicate):
    column: int
    value: ScalarValue


class Eq(_Cmp):
    pass


class Ne(_Cmp):
    pass


class Lt(_Cmp):
    pass


class Le(_Cmp):
    pass


class Gt(_Cmp):
    pass


class Ge(_Cmp):
    pass


@dataclass
class IsNull(Predicate):
    column: int


@dataclass
class IsNotNull(Predicate):
    column: int


@dataclass
class And(Predicate):
    left: Predicate
    right: Predicate


@dataclass
class Or(Predicate):
    left: Predicate
    right: Predicate


@dataclass
class Not(Predicate):
    inner: Predicate


@dataclass
class Like(Predicate):
    column: int
    pattern: str


@dataclass
class ILike(Predicate):
    column: int
    pattern: str


# ----------------------------------------------------------------- ScalarExpr
class ArithmeticOp(IntEnum):
    """ArithmeticOp (filter_project.rs:185-192)."""
    Add = 0
    Subtract = 1
    Multiply = 2
    Divide = 3


class ScalarExpr:
    """ScalarExpr (filter_project.rs:194-213)."""


@dataclass
class ColumnRef(ScalarExpr):
    index: int


@dataclass
class LiteralExpr(ScalarExpr):
    value: ScalarValue


@dataclass
class BinaryOp(ScalarExpr):
    op: ArithmeticOp
    left: ScalarExpr
    right: ScalarExpr


@dataclass
class Negate(ScalarExpr):
    inner: ScalarExpr


@dataclass
class IfThen(ScalarExpr):
    clauses: List[Tuple[Predicate, ScalarExpr]]
    else_expr: Optional[ScalarExpr] = None


# ------------------------------------------------------------ node flattening
class _Nodes:
    """Predicate and expression trees as the C ABI's node arrays."""

    def __init__(self):
        self.preds: List[L.PredNode] = []
        self.exprs: List[L.ExprNode] = []
        self.keep = []  # bytes and int arrays referenced by the nodes

    def _scalar(self, v: ScalarValue) -> L.Scalar:
        s = L.Scalar()
        s.type = _SCALAR_CODE[v.kind]
        if v.kind == "Float64":
            s.f64 = float(v.value)
        elif v.kind == "Utf8":
            b = str(v.value).encode()
            self.keep.append(b)
            s.str, s.str_len = b, len(b)
        else:
            s.i64 = int(v.value)
        return s

    def pred(self, p: Predicate) -> int:
        i = len(self.preds)
        n = L.PredNode()
        self.preds.append(n)
        if isinstance(p, Literal):
            n.kind = L.PRED["LITERAL"]
            n.value.type = _SCALAR_CODE["Boolean"]
            n.value.i64 = 1 if p.value else 0
        elif isinstance(p, _Cmp):
            n.kind = L.PRED[{Eq: "EQ", Ne: "NE", Lt: "LT", Le: "LE", Gt: "GT", Ge: "GE"}[type(p)]]
            n.column = p.column
            n.value = self._scalar(p.value)
        elif isinstance(p, (IsNull, IsNotNull)):
            n.kind = L.PRED["IS_NULL" if isinstance(p, IsNull) else "IS_NOT_NULL"]
            n.column = p.column
        elif isinstance(p, (And, Or)):
            n.kind = L.PRED["AND" if isinstance(p, And) else "OR"]
            n.left = self.pred(p.left)
            n.right = self.pred(p.right)
        elif isinstance(p, Not):
            n.kind = L.PRED["NOT"]
            n.left = self.pred(p.inner)
        elif isinstance(p, (Like, ILike)):
            n.kind = L.PRED["LIKE" if isinstance(p, Like) else "ILIKE"]
            n.column = p.column
            b = p.pattern.encode()
            self.keep.append(b)
            n.pattern, n.pattern_len = b, len(b)
        else:
            raise TypeError(f"not a Predicate: {p!r}")
        return i

    def expr(self, e: ScalarExpr) -> int:
        i = len(self.exprs)
        n = L.ExprNode()
        n.else_expr = -1
        self.exprs.append(n)
        if isinstance(e, ColumnRef):
            n.kind = L.EXPR["COLUMN"]
            n.column = e.index
        elif isinstance(e, LiteralExpr):
            n.kind = L.EXPR["LITERAL"]
            n.value = self._scalar(e.value)
        elif isinstance(e, BinaryOp):
            n.kind = L.EXPR["BINARY"]
            n.op = int(e.op)
            n.left = self.expr(e.left)
            n.right = self.expr(e.right)
        elif isinstance(e, Negate):
            n.kind = L.EXPR["NEGATE"]
            n.left = self.expr(e.inner)
        elif isinstance(e, IfThen):
            n.kind = L.EXPR["IF_THEN"]
            ps = [self.pred(p) for p, _ in e.clauses]
            xs = [self.expr(x) for _, x in e.clauses]
            cp = (ctypes.c_int32 * max(len(ps), 1))(*ps)
            cx = (ctypes.c_int32 * max(len(xs), 1))(*xs)
            self.keep += [cp, cx]
            n.clause_preds = ctypes.cast(cp, ctypes.POINTER(ctypes.c_int32))
            n.clause_exprs = ctypes.cast(cx, ctypes.POINTER(ctypes.c_int32))
            n.n_clauses = len(ps)
            n.else_expr = self.expr(e.else_expr) if e.else_expr is not None else -1
        else:
            raise TypeError(f"not a ScalarExpr: {e!r}")
        return i

    def pred_array(self):
        return (L.PredNode * max(len(self.preds), 1))(*self.preds), len(self.preds)

    def expr_array(self):
        return (L.ExprNode * max(len(self.exprs), 1))(*self.exprs), len(self.exprs)


def _pred_nodes(predicate: Predicate):
    n = _Nodes()
    n.pred(predicate)
    arr, count = n.pred_array()
    return arr, count, n


# ------------------------------------------------------------------ operators
def evaluate_predicate(batch: pa.RecordBatch, predicate: Predicate) -> pa.BooleanArray:
    """evaluate_predicate (filter_project.rs:64-123)."""
    ctx = context()
    e = Exported(batch)
    nodes, count, keep = _pred_nodes(predicate)
    oa, os_ = L.ArrowArray(), L.ArrowSchema()
    L.check(ctx.lib.prism_evaluate_predicate(ctx.handle, e.aptr, e.sptr, nodes, count, ctypes.addressof(oa),
                                             ctypes.addressof(os_)))
    return import_array(oa, os_)


def filter_batch(batch: pa.RecordBatch, mask: pa.BooleanArray) -> pa.RecordBatch:
    """filter_batch (filter_project.rs:125-135)."""
    ctx = context()
    e = Exported(batch)
    ma, ms = L.ArrowArray(), L.ArrowSchema()
    mask._export_to_c(ctypes.addressof(ma), ctypes.addressof(ms))
    try:
        oa, os_ = L.ArrowArray(), L.ArrowSchema()
        L.check(ctx.lib.prism_filter_batch(ctx.handle, e.aptr, e.sptr, ctypes.addressof(ma), ctypes.addressof(ms),
                                           ctypes.addressof(oa), ctypes.addressof(os_)))
        return import_batch(oa, os_)
    finally:
        for s in (ma, ms):
            if s.release:
                ctypes.CFUNCTYPE(None, ctypes.c_void_p)(s.release)(ctypes.addressof(s))


def project_batch(batch: pa.RecordBatch, column_indices: Sequence[int]) -> pa.RecordBatch:
    """project_batch (filter_project.rs:137-160)."""
    ctx = context()
    e = Exported(batch)
    idx = u32s(column_indices)
    oa, os_ = L.ArrowArray(), L.ArrowSchema()
    L.check(ctx.lib.prism_project_batch(ctx.handle, e.aptr, e.sptr, idx, len(column_indices), ctypes.addressof(oa),
                                        ctypes.addressof(os_)))
    return import_batch(oa, os_)


def filter_and_project(batch: pa.RecordBatch, predicate: Optional[Predicate],
                       column_indices: Sequence[int]) -> pa.RecordBatch:
    """filter_and_project (filter_project.rs:162-180)."""
    ctx = context()
    e = Exported(batch)
    idx = u32s(column_indices)
    if predicate is None:
        nodes, count, keep = None, 0, None
    else:
        nodes, count, keep = _pred_nodes(predicate)
    oa, os_ = L.ArrowArray(), L.ArrowSchema()
    L.check(ctx.lib.prism_filter_and_project(ctx.handle, e.aptr, e.sptr, nodes, count, idx, len(column_indices),
                                             ctypes.addressof(oa), ctypes.addressof(os_)))
    return import_batch(oa, os_)


def evaluate_scalar_expr(batch: pa.RecordBatch, expr: ScalarExpr) -> pa.Array:
    """evaluate_scalar_expr (filter_project.rs:215-294)."""
    ctx = context()
    e = Exported(batch)
    n = _Nodes()
    root = n.expr(expr)
    ex, nex = n.expr_array()
    pr, npr = n.pred_array()
    oa, os_ = L.ArrowArray(), L.ArrowSchema()
    L.check(ctx.lib.prism_evaluate_scalar_expr(ctx.handle, e.aptr, e.sptr, ex, nex, pr, npr, root,
                                               ctypes.addressof(oa), ctypes.addressof(os_)))
    return import_array(oa, os_)


def project_batch_with_exprs(batch: pa.RecordBatch, column_indices: Sequence[int],
                             expressions: Sequence[ScalarExpr]) -> pa.RecordBatch:
    """project_batch_with_exprs (filter_project.rs:304-333)."""
    ctx = context()
    e = Exported(batch)
    n = _Nodes()
    roots = [n.expr(x) for x in expressions]
    ex, nex = n.expr_array()
    pr, npr = n.pred_array()
    idx = u32s(column_indices)
    rt = (ctypes.c_int32 * max(len(roots), 1))(*roots)
    oa, os_ = L.ArrowArray(), L.ArrowSchema()
    L.check(ctx.lib.prism_project_batch_with_exprs(ctx.handle, e.aptr, e.sptr, idx, len(column_indices), ex, nex, rt,
                                                   len(roots), pr, npr, ctypes.addressof(oa), ctypes.addressof(os_)))
    return import_batch(oa, os_)
