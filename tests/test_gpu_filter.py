"""filter_project / string_ops on the GPU against the oracle
(filter_project.rs:20-404, string_ops.rs:16-25), the reference's own tests
first, then seeded random batches with nulls, sliced inputs and edge cases."""
import json
import os

import numpy as np
import pyarrow as pa
import pytest

from gpu_util import say

import filter_oracle as fo
from prism_executor import _lib as L
from prism_executor import filter_project as fp
from prism_executor import string_ops as so

pytestmark = pytest.mark.gpu
SV = fp.ScalarValue
KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def same(a: pa.RecordBatch, b: pa.RecordBatch) -> bool:
    """Batch equality with NaN equal to NaN (RecordBatch.equals says NaN != NaN)."""
    if a.schema.names != b.schema.names or a.num_rows != b.num_rows:
        return False
    for x, y in zip(a.columns, b.columns):
        if x.type != y.type or not x.is_valid().equals(y.is_valid()):
            return False
        if pa.types.is_floating(x.type):
            if not np.array_equal(np.asarray(x.fill_null(0)), np.asarray(y.fill_null(0)), equal_nan=True):
                return False
        elif x.to_pylist() != y.to_pylist():
            return False
    return True


def kat_batch():
    b = KATS["filter_project"]["batch"]
    return pa.record_batch([pa.array(b["id"], pa.int64()), pa.array(b["name"]), pa.array(b["value"])],
                           names=["id", "name", "value"])


def test_reference_tests(ctx):
    say("filter_project reference tests")
    b = kat_batch()
    mask = fp.evaluate_predicate(b, fp.Gt(0, SV.Int64(3)))
    r = fp.filter_batch(b, mask)
    assert r.num_rows == 2 and r.column(0).to_pylist() == [4, 5]
    r = fp.filter_and_project(b, fp.Le(2, SV.Float64(30.0)), [1, 2])
    assert (r.num_rows, r.num_columns) == (3, 2) and r.column(0)[0].as_py() == "alice" and r.column(0)[2].as_py() == "carol"
    mask = fp.evaluate_predicate(b, fp.And(fp.Ge(0, SV.Int64(2)), fp.Le(0, SV.Int64(4))))
    assert fp.filter_batch(b, mask).num_rows == 3
    mask = fp.evaluate_predicate(b, fp.Literal(True))
    assert len(mask) == 5 and all(v is True for v in mask.to_pylist())
    b2 = pa.record_batch([pa.array([1, 2, 3], pa.int64()), pa.array(["a", "b", "c"])], names=["id", "label"])
    out = fp.project_batch_with_exprs(b2, [0], [fp.IfThen([(fp.Eq(0, SV.Int64(1)), fp.LiteralExpr(SV.Utf8("owner")))],
                                                          fp.ColumnRef(1))])
    assert out.column(1).to_pylist() == ["owner", "b", "c"] and out.schema.names == ["id", "expr_0"]
    assert so.string_like(pa.array(["hello", "world", "help", "held"]), "hel%").to_pylist() == [True, False, True, True]
    assert so.string_like(pa.array(["Hello World", "Goodbye"]), "%World%").to_pylist() == [True, False]


def _rand_batch(n, seed):
    rng = np.random.default_rng(seed)
    nul = lambda: rng.random(n) < 0.1  # noqa: E731
    words = ["", "a", "abc", "xabcx", "zz", "Hello", "hello", "été", "ab\\c", "a%b"]
    f = rng.standard_normal(n) * 10
    f[::97] = np.nan
    f[::89] = -0.0
    f[::83] = 0.0
    return pa.record_batch([
        pa.array(rng.integers(-50, 50, n).astype(np.int32), mask=nul()),
        pa.array(rng.integers(-10**12, 10**12, n), mask=nul()),
        pa.array(f, mask=nul()),
        pa.array([words[i] for i in rng.integers(0, len(words), n)], mask=nul()),
        pa.array(rng.integers(0, 3000, n).astype(np.int32), pa.int32(), mask=nul()).cast(pa.date32()),
        pa.array(rng.random(n) < 0.5, mask=nul()),
    ], names=["i32", "i64", "f64", "s", "d", "b"])


PREDS = [
    fp.Gt(0, SV.Int32(7)), fp.Le(1, SV.Int64(0)), fp.Eq(2, SV.Float64(float("nan"))), fp.Lt(2, SV.Float64(0.0)),
    fp.Eq(2, SV.Float64(-0.0)), fp.Ge(2, SV.Float64(0.0)), fp.Ne(3, SV.Utf8("abc")), fp.Lt(3, SV.Utf8("b")),
    fp.Ge(4, SV.Date32(1500)), fp.Lt(4, SV.Int32(10)), fp.IsNull(3), fp.IsNotNull(5),
    fp.And(fp.Gt(0, SV.Int32(0)), fp.Lt(2, SV.Float64(1.0))),
    fp.Or(fp.Not(fp.Eq(3, SV.Utf8("zz"))), fp.IsNull(0)),
    fp.Like(3, "%abc%"), fp.Like(3, "_"), fp.Like(3, "h_llo"), fp.Like(3, "%"), fp.Like(3, "ab\\\\c"),
    fp.Like(3, "a\\%b"), fp.ILike(3, "HELLO"), fp.ILike(3, "%ABC%"), fp.Literal(False),
    fp.Not(fp.And(fp.Or(fp.Gt(0, SV.Int32(-20)), fp.Lt(1, SV.Int64(5))), fp.Not(fp.IsNull(2)))),
    fp.Like(3, "he%"), fp.Like(3, "%bc"), fp.Like(3, "zz"), fp.Like(3, "%%"), fp.Like(3, ""), fp.ILike(3, "%é%"),
    fp.ILike(3, "HE%"), fp.ILike(3, "%LO"),
]


@pytest.mark.parametrize("k", range(len(PREDS)))
def test_predicates_match_oracle(ctx, k):
    say(f"predicate {k}: {PREDS[k]}")
    b = _rand_batch(20_000, 100 + k)
    v, m = fo.eval_pred(b, PREDS[k])
    got = fp.evaluate_predicate(b, PREDS[k])
    assert got.to_pylist() == fo.mask_array(v, m).to_pylist()
    assert got.null_count == int((~m).sum())
    r = fp.filter_batch(b, got)
    assert same(r, fo.filter_rows(b, v, m))


def test_sliced_inputs_and_nullable_mask(ctx):
    say("sliced batch + nullable mask")
    b = _rand_batch(10_000, 7).slice(1234, 5000)
    v, m = fo.eval_pred(b, fp.Gt(0, SV.Int32(5)))
    mask = fo.mask_array(v, m)
    assert same(fp.filter_batch(b, mask), fo.filter_rows(b, v, m))
    sm = pa.array(np.random.default_rng(1).random(12_000) < 0.3).slice(3, 5000)
    sv, smm = np.asarray(sm.to_pylist(), bool), np.ones(5000, bool)
    assert same(fp.filter_batch(b, sm), fo.filter_rows(b, sv, smm))
    got = fp.evaluate_predicate(b, fp.Like(3, "%abc%"))
    v, m = fo.eval_pred(b, fp.Like(3, "%abc%"))
    assert got.to_pylist() == fo.mask_array(v, m).to_pylist()


def test_project_and_filter_and_project(ctx):
    say("project / filter_and_project")
    b = _rand_batch(8_000, 9)
    p = fp.project_batch(b, [3, 0])
    assert same(p, b.select([3, 0]))
    e = fp.project_batch(b, [])
    assert e.num_columns == 0 and e.num_rows == b.num_rows
    pred = fp.And(fp.Gt(1, SV.Int64(0)), fp.Like(3, "%a%"))
    v, m = fo.eval_pred(b, pred)
    assert same(fp.filter_and_project(b, pred, [5, 3, 2]), fo.filter_rows(b, v, m).select([5, 3, 2]))
    assert same(fp.filter_and_project(b, pred, []), fo.filter_rows(b, v, m))
    assert same(fp.filter_and_project(b, None, [1]), b.select([1]))
    none = fp.filter_and_project(b, fp.Literal(False), [0, 3])
    assert none.num_rows == 0 and none.schema.names == ["i32", "s"]


EXPRS = [
    fp.BinaryOp(fp.ArithmeticOp.Multiply, fp.ColumnRef(1),
                fp.BinaryOp(fp.ArithmeticOp.Subtract, fp.LiteralExpr(SV.Int64(100)), fp.ColumnRef(0))),
    fp.BinaryOp(fp.ArithmeticOp.Divide, fp.ColumnRef(2), fp.ColumnRef(0)),
    fp.BinaryOp(fp.ArithmeticOp.Add, fp.ColumnRef(5), fp.LiteralExpr(SV.Float64(0.5))),
    fp.Negate(fp.ColumnRef(2)),
    fp.IfThen([(fp.Gt(0, SV.Int32(10)), fp.ColumnRef(2)), (fp.Lt(0, SV.Int32(-10)), fp.LiteralExpr(SV.Float64(-1)))],
              fp.LiteralExpr(SV.Float64(0.5))),
    fp.IfThen([(fp.IsNull(2), fp.ColumnRef(1)), (fp.Gt(2, SV.Float64(0.0)), fp.ColumnRef(0))]),
    fp.IfThen([(fp.Like(3, "%abc%"), fp.LiteralExpr(SV.Utf8("has-abc"))), (fp.Eq(0, SV.Int32(3)), fp.ColumnRef(3))],
              fp.LiteralExpr(SV.Utf8(""))),
    fp.IfThen([(fp.Gt(1, SV.Int64(0)), fp.ColumnRef(3))]),
    fp.ColumnRef(4),
    fp.LiteralExpr(SV.Int32(42)),
    fp.IfThen([], fp.ColumnRef(3)),
]


@pytest.mark.parametrize("k", range(len(EXPRS)))
def test_scalar_exprs_match_oracle(ctx, k):
    say(f"expression {k}")
    b = _rand_batch(6_000, 200 + k)
    ref = fo.eval_expr(b, EXPRS[k])
    got = fp.evaluate_scalar_expr(b, EXPRS[k])
    assert got.type == ref.type
    if pa.types.is_floating(ref.type):
        assert got.is_valid().equals(ref.is_valid())
        assert np.array_equal(np.asarray(got.fill_null(0)), np.asarray(ref.fill_null(0)), equal_nan=True)
    else:
        assert got.to_pylist() == ref.to_pylist()


def test_project_batch_with_exprs_schema(ctx):
    say("project_batch_with_exprs")
    b = _rand_batch(3_000, 11)
    exprs = [EXPRS[0], EXPRS[6]]
    got = fp.project_batch_with_exprs(b, [3, 1], exprs)
    ref = fo.project_with_exprs(b, [3, 1], exprs)
    assert got.schema.names == ref.schema.names == ["s", "i64", "expr_0", "expr_1"]
    for i in range(4):
        assert got.column(i).to_pylist() == ref.column(i).to_pylist()


def test_string_ilike_and_nulls(ctx):
    say("string_ilike")
    arr = pa.array(["Hello", None, "HELLO world", "help", "", None])
    assert so.string_ilike(arr, "hello%").to_pylist() == [True, None, True, False, False, None]
    assert so.string_like(arr, "%").to_pylist() == [True, None, True, True, True, None]


def test_type_errors_match_reference_messages(ctx):
    say("type errors")
    b = kat_batch()
    with pytest.raises(L.PrismError) as e:
        fp.evaluate_predicate(b, fp.Eq(0, SV.Int32(1)))
    assert e.value.kind == "InvalidArgument" and "Cannot compare column type Int64 with scalar Int32(1)" in str(e.value)
    with pytest.raises(L.PrismError) as e:
        fp.evaluate_predicate(b, fp.Like(0, "x%"))
    assert e.value.kind == "InvalidArgument" and "LIKE requires Utf8 column, got Int64" in str(e.value)
    with pytest.raises(L.PrismError) as e:
        fp.evaluate_scalar_expr(b, fp.IfThen([]))
    assert e.value.kind == "InvalidArgument"


def test_empty_batch(ctx):
    say("empty batch")
    b = _rand_batch(10, 1).slice(0, 0)
    assert len(fp.evaluate_predicate(b, fp.Gt(0, SV.Int32(0)))) == 0
    assert fp.filter_and_project(b, fp.Gt(0, SV.Int32(0)), [0, 3]).num_rows == 0
    assert len(fp.evaluate_scalar_expr(b, EXPRS[0])) == 0
    assert len(fp.evaluate_scalar_expr(b, EXPRS[6])) == 0


@pytest.mark.parametrize("pattern,ci", [("%abc%", False), ("%ab%", False), ("abc%", False), ("%abc", False),
                                        ("%a%", False), ("%xyzw%", False), ("%q%z%", False), ("%ABC%", True),
                                        ("aB%", True)])
def test_like_shapes_on_long_strings(ctx, pattern, ci):
    say(f"like shapes {pattern} ci={ci}")
    from prism_executor import datagen as dg
    b, o = dg.lower_utf8(60_001, 11, 3, 1, 64)
    strs = pa.StringArray.from_buffers(60_001, pa.py_buffer(o.tobytes()), pa.py_buffer(b.tobytes()))
    for arr in (strs, strs.slice(3, 40_000), strs.slice(1, 17)):
        v, m = fo.eval_pred(pa.record_batch([arr], names=["s"]), (fp.ILike if ci else fp.Like)(0, pattern))
        got = (so.string_ilike if ci else so.string_like)(arr, pattern)
        assert got.to_pylist() == fo.mask_array(v, m).to_pylist()


@pytest.mark.parametrize("ci", [False, True])
def test_like_contains_random_literals(ctx, ci):
    """CONTAINS patterns of 1..12 literal bytes over strings of 0..100 bytes
    from a small alphabet (frequent partial matches, letters of both cases
    beside the bytes that differ from a letter by 0x20, a two-byte UTF-8
    character), every alignment of string starts, matches across 16-byte
    blocks and at both string ends, against the filter oracle."""
    say(f"like contains random literals ci={ci}")
    rng = np.random.default_rng(29 + ci)
    alpha = ["a", "b", "A", "B", "@", "[", "`", "{", "\u00e9"]
    strs = ["".join(rng.choice(alpha, int(rng.integers(0, 101)))) for _ in range(6000)]
    arr = pa.array(strs, pa.string())
    for lit_len in range(1, 13):
        for _ in range(2):
            lit = "".join(rng.choice(alpha[:8], lit_len))
            pattern = "%" + lit + "%"
            for a in (arr, arr.slice(5, 3001)):
                v, m = fo.eval_pred(pa.record_batch([a], names=["s"]), (fp.ILike if ci else fp.Like)(0, pattern))
                got = (so.string_ilike if ci else so.string_like)(a, pattern)
                assert got.to_pylist() == fo.mask_array(v, m).to_pylist(), pattern
