"""Pins of the filter/project oracle (oracle/filter_oracle.py): the reference's
own known answers (filter_project.rs:434-509, string_ops.rs:125-164), IEEE
totalOrder from its definition, and Arrow C++ (pyarrow) kernels where arrow-rs
and Arrow C++ agree (no NaN or signed zero in the data)."""
import json
import os

import numpy as np
import pyarrow as pa
import pyarrow.compute as pc
import pytest

import filter_oracle as fo
from prism_executor import filter_project as fp

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
SV = fp.ScalarValue


def kat_batch():
    b = KATS["filter_project"]["batch"]
    return pa.record_batch([pa.array(b["id"], pa.int64()), pa.array(b["name"]), pa.array(b["value"])],
                           names=["id", "name", "value"])


def test_reference_kats():
    k = KATS["filter_project"]
    b = kat_batch()
    v, m = fo.eval_pred(b, fp.Gt(0, SV.Int64(3)))
    assert fo.filter_rows(b, v, m).column(0).to_pylist() == k["gt_id_3"]["ids"]
    v, m = fo.eval_pred(b, fp.Le(2, SV.Float64(30.0)))
    r = fo.filter_rows(b, v, m).select([1, 2])
    assert (r.num_rows, r.num_columns) == (k["le_value_30_project_1_2"]["rows"], k["le_value_30_project_1_2"]["columns"])
    assert r.column(0)[0].as_py() == k["le_value_30_project_1_2"]["first"]
    assert r.column(0)[2].as_py() == k["le_value_30_project_1_2"]["third"]
    v, m = fo.eval_pred(b, fp.And(fp.Ge(0, SV.Int64(2)), fp.Le(0, SV.Int64(4))))
    assert fo.filter_rows(b, v, m).num_rows == k["ge2_and_le4"]["rows"]
    v, m = fo.eval_pred(b, fp.Literal(True))
    assert bool((v & m).all()) == k["literal_true"]["all_true"]
    t = k["if_then"]
    b2 = pa.record_batch([pa.array(t["id"], pa.int64()), pa.array(t["label"])], names=["id", "label"])
    out = fo.project_with_exprs(b2, [0], [fp.IfThen([(fp.Eq(0, SV.Int64(1)), fp.LiteralExpr(SV.Utf8("owner")))],
                                                    fp.ColumnRef(1))])
    assert out.column(1).to_pylist() == t["expected"] and out.schema.names == ["id", "expr_0"]
    for name in ("like_hel", "contains_world"):
        s = KATS["string_ops"][name]
        b3 = pa.record_batch([pa.array(s["values"])], names=["s"])
        v, m = fo.eval_pred(b3, fp.Like(0, s["pattern"]))
        assert (v & m).tolist() == s["expected"]


def test_total_order_from_definition():
    # IEEE 754 totalOrder: -NaN < -inf < -1.5 < -0.0 < +0.0 < 1.5 < inf < +NaN
    neg_nan = -np.float64(np.nan) if np.signbit(-np.float64(np.nan)) else np.float64(np.nan)
    vals = np.array([np.nan, 1.5, -0.0, np.inf, 0.0, -np.inf, -1.5, neg_nan])
    order = [7, 5, 6, 2, 4, 1, 3, 0]
    keys = fo.total_key(vals)
    assert np.argsort(keys, kind="stable").tolist() == order
    b = pa.record_batch([pa.array(vals)], names=["x"])
    v, _ = fo.eval_pred(b, fp.Eq(0, SV.Float64(float("nan"))))
    assert v.tolist() == [True, False, False, False, False, False, False, False]
    v, _ = fo.eval_pred(b, fp.Lt(0, SV.Float64(0.0)))
    assert v.tolist() == [False, False, True, False, False, True, True, True]
    v, _ = fo.eval_pred(b, fp.Eq(0, SV.Float64(-0.0)))
    assert v.tolist() == [False, False, True, False, False, False, False, False]


def _rand_batch(n, seed):
    rng = np.random.default_rng(seed)
    nul = lambda: rng.random(n) < 0.1  # noqa: E731
    words = ["", "a", "abc", "xabcx", "zz", "Hello", "hello", "été"]
    return pa.record_batch([
        pa.array(rng.integers(-50, 50, n).astype(np.int32), mask=nul()),
        pa.array(rng.integers(-10**12, 10**12, n), mask=nul()),
        pa.array(rng.standard_normal(n) * 10, mask=nul()),
        pa.array([words[i] for i in rng.integers(0, len(words), n)], mask=nul()),
        pa.array(rng.integers(0, 3000, n).astype(np.int32), pa.int32(), mask=nul()).cast(pa.date32()),
    ], names=["i32", "i64", "f64", "s", "d"])


CMPS = [(fp.Eq, pc.equal), (fp.Ne, pc.not_equal), (fp.Lt, pc.less), (fp.Le, pc.less_equal), (fp.Gt, pc.greater),
        (fp.Ge, pc.greater_equal)]


@pytest.mark.parametrize("cls,fn", CMPS)
def test_comparisons_against_arrow_cpp(cls, fn):
    b = _rand_batch(5000, 3)
    cases = [(0, SV.Int32(7), pa.scalar(7, pa.int32())), (1, SV.Int64(123), pa.scalar(123, pa.int64())),
             (2, SV.Float64(0.25), pa.scalar(0.25)), (3, SV.Utf8("abc"), pa.scalar("abc")),
             (4, SV.Int32(1500), pa.scalar(1500, pa.int32()).cast(pa.date32()))]
    for col, sv, ps in cases:
        v, m = fo.eval_pred(b, cls(col, sv))
        ref = fn(b.column(col), ps)
        assert m.tolist() == ref.is_valid().to_pylist()
        assert [bool(x) for x, ok in zip(v, m) if ok] == [x for x in ref.to_pylist() if x is not None]


def test_boolean_logic_null_propagation_against_arrow_cpp():
    b = _rand_batch(4000, 4)
    p1, p2 = fp.Gt(0, SV.Int32(0)), fp.Lt(2, SV.Float64(1.0))
    for cls, fn in ((fp.And, pc.and_), (fp.Or, pc.or_)):
        v, m = fo.eval_pred(b, cls(p1, p2))
        ref = fn(pc.greater(b.column(0), 0), pc.less(b.column(2), 1.0))
        assert m.tolist() == ref.is_valid().to_pylist()
        assert [bool(x) for x, ok in zip(v, m) if ok] == [x for x in ref.to_pylist() if x is not None]
    v, m = fo.eval_pred(b, fp.Not(p1))
    ref = pc.invert(pc.greater(b.column(0), 0))
    assert [bool(x) if ok else None for x, ok in zip(v, m)] == ref.to_pylist()
    v, m = fo.eval_pred(b, fp.IsNull(3))
    assert v.tolist() == pc.is_null(b.column(3)).to_pylist() and m.all()


def test_like_against_arrow_cpp():
    b = _rand_batch(3000, 5)
    for pat in ("%abc%", "a%", "%c", "_", "__", "h_llo", "%", "", "\\%x"):
        v, m = fo.eval_pred(b, fp.Like(3, pat))
        ref = pc.match_like(b.column(3), pat)
        assert [bool(x) if ok else None for x, ok in zip(v, m)] == ref.to_pylist(), pat
    v, m = fo.eval_pred(b, fp.ILike(3, "HELLO"))
    ref = pc.match_like(b.column(3), "HELLO", ignore_case=True)
    assert [bool(x) if ok else None for x, ok in zip(v, m)] == ref.to_pylist()


def test_filter_against_arrow_cpp():
    b = _rand_batch(2000, 6)
    v, m = fo.eval_pred(b, fp.Gt(1, SV.Int64(0)))
    assert fo.filter_rows(b, v, m).equals(b.filter(fo.mask_array(v, m)))


def test_arithmetic_and_if_then_against_arrow_cpp():
    b = _rand_batch(3000, 7)
    f = lambda i: b.column(i).cast(pa.float64())  # noqa: E731
    e = fp.BinaryOp(fp.ArithmeticOp.Multiply, fp.ColumnRef(1),
                    fp.BinaryOp(fp.ArithmeticOp.Subtract, fp.LiteralExpr(SV.Int64(100)), fp.ColumnRef(0)))
    assert fo.eval_expr(b, e).equals(pc.multiply(f(1), pc.subtract(pa.scalar(100.0), f(0))))
    e = fp.BinaryOp(fp.ArithmeticOp.Divide, fp.ColumnRef(2), fp.ColumnRef(0))
    ref = pc.divide(f(2), f(0))
    got = fo.eval_expr(b, e)
    assert got.is_valid().equals(ref.is_valid())
    assert np.array_equal(np.asarray(got.fill_null(0)), np.asarray(ref.fill_null(0)), equal_nan=True)
    assert fo.eval_expr(b, fp.Negate(fp.ColumnRef(0))).equals(pc.negate(f(0)))
    e = fp.IfThen([(fp.Gt(0, SV.Int32(10)), fp.ColumnRef(2)), (fp.Lt(0, SV.Int32(-10)), fp.LiteralExpr(SV.Float64(-1)))],
                  fp.LiteralExpr(SV.Float64(0.5)))
    c1 = pc.fill_null(pc.greater(b.column(0), 10), False)
    c2 = pc.fill_null(pc.less(b.column(0), -10), False)
    ref = pc.if_else(c1, b.column(2), pc.if_else(c2, pa.scalar(-1.0), pa.scalar(0.5)))
    assert fo.eval_expr(b, e).equals(ref)


def test_type_errors():
    b = kat_batch()
    with pytest.raises(fo.OracleError) as e:
        fo.eval_pred(b, fp.Eq(0, SV.Int32(1)))
    assert e.value.kind == "InvalidArgument" and "Cannot compare column type Int64" in str(e.value)
    with pytest.raises(fo.OracleError) as e:
        fo.eval_pred(b, fp.Like(0, "x%"))
    assert "LIKE requires Utf8 column, got Int64" in str(e.value)
