"""Parquet footer, statistics and row-group skipping (parquet_scan.rs:55-494)
on the CPU: the library's footer parser against pyarrow's reader of the same
files, and the scan's skipping rule against the oracle restatement, pinned by
the reference's own test expectations (parquet_scan.rs:560-720: 500 rows in 5
row groups, Lt(value, 150.0) keeps 200 rows, Eq(name, "gamma") keeps 100) and
by soundness on the data (a skipped row group holds no matching row)."""
import io
import os

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

import filter_oracle as fo
import parquet_oracle as po
from prism_executor import _lib as L
from prism_executor.filter_project import (And, Eq, Ge, Gt, IsNull, Le, Like, Literal, Lt, Ne, Not, Or,
                                           ScalarValue)
from prism_executor.parquet_scan import (ParquetFile, list_files, parquet_row_count, row_group_might_match,
                                         stats_might_match, ColumnStats)


def write(table, **opts):
    buf = io.BytesIO()
    pq.write_table(table, buf, **opts)
    return buf.getvalue()


def reference_table():
    """build_test_parquet_bytes (parquet_scan.rs:560-598): id 0..499, value = id,
    name alpha/beta/gamma/delta/epsilon by hundreds; fields not nullable."""
    ids = np.arange(500, dtype=np.int64)
    names = [["alpha", "beta", "gamma", "delta", "epsilon"][i // 100] for i in range(500)]
    schema = pa.schema([pa.field("id", pa.int64(), False), pa.field("value", pa.float64(), False),
                        pa.field("name", pa.utf8(), False)])
    return pa.table([pa.array(ids), pa.array(ids.astype(np.float64)), pa.array(names)], schema=schema)


def reference_bytes():
    return write(reference_table(), row_group_size=100, compression="snappy", write_statistics=True)


def kept_rows(f, pred):
    return sum(f.row_group_rows(g) for g in range(f.num_row_groups) if row_group_might_match(pred, f, g))


def test_reference_fixture_metadata():
    f = ParquetFile(reference_bytes())
    assert f.num_row_groups == 5 and f.num_columns == 3 and f.num_rows == 500
    assert [(n, t, nl) for n, t, nl, _ in f.fields] == [("id", 2, False), ("value", 3, False), ("name", 4, False)]
    assert [f.row_group_rows(g) for g in range(5)] == [100] * 5
    s = f.statistics(2, 2)
    assert (s.min, s.max) == (b"gamma", b"gamma")
    s = f.statistics(4, 1)
    assert (s.min, s.max, s.null_count) == (400.0, 499.0, 0)


def test_reference_skipping_expectations():
    f = ParquetFile(reference_bytes())
    assert kept_rows(f, Lt(1, ScalarValue.Float64(150.0))) == 200      # test_row_group_skipping
    assert kept_rows(f, Eq(2, ScalarValue.Utf8("gamma"))) == 100        # test_string_stats_skipping
    assert kept_rows(f, Literal(True)) == 500


def test_row_count_over_a_directory(tmp_path):
    d = tmp_path / "t"
    (d / "sub").mkdir(parents=True)
    (d / "sub" / "test.parquet").write_bytes(reference_bytes())
    (d / "notes.txt").write_text("not parquet")
    assert list_files(str(d)) == [str(d / "sub" / "test.parquet")]
    assert parquet_row_count([str(d)]) == 500                           # test_row_count_metadata
    assert parquet_row_count(["file://" + str(d)], Lt(1, ScalarValue.Float64(150.0))) == 200


def random_table(n, seed):
    rng = np.random.default_rng(seed)
    strs = np.array(["".join(chr(97 + c) for c in rng.integers(0, 26, rng.integers(1, 6))) for _ in range(64)])
    return pa.table({
        "i32": pa.array(rng.integers(-1000, 1000, n).astype(np.int32), mask=rng.random(n) < 0.1),
        "i64": pa.array(np.sort(rng.integers(-10**6, 10**6, n))),
        "f64": pa.array(rng.standard_normal(n) * 100, mask=rng.random(n) < 0.05),
        "s": pa.array(strs[rng.integers(0, 64, n)], mask=rng.random(n) < 0.05),
        "d": pa.array(rng.integers(0, 2557, n).astype(np.int32)).cast(pa.date32()),
        "b": pa.array(rng.random(n) < 0.5),
    })


def random_predicates(rng, k):
    scal = [lambda: ScalarValue.Int32(int(rng.integers(-1100, 1100))),
            lambda: ScalarValue.Int64(int(rng.integers(-1_100_000, 1_100_000))),
            lambda: ScalarValue.Float64(float(rng.standard_normal() * 150)),
            lambda: ScalarValue.Utf8("".join(chr(97 + c) for c in rng.integers(0, 26, rng.integers(1, 4)))),
            lambda: ScalarValue.Date32(int(rng.integers(0, 2557)))]
    col_for = {"Int32": [0, 4], "Int64": [1], "Float64": [2], "Utf8": [3], "Date32": [4]}

    def leaf():
        v = scal[int(rng.integers(0, len(scal)))]()
        cols = col_for[v.kind]
        c = cols[int(rng.integers(0, len(cols)))]
        cls = [Eq, Lt, Le, Gt, Ge, Ne][int(rng.integers(0, 6))]
        return cls(c, v)

    def tree(d):
        r = rng.random()
        if d == 0 or r < 0.5:
            return leaf()
        if r < 0.7:
            return And(tree(d - 1), tree(d - 1))
        if r < 0.85:
            return Or(tree(d - 1), tree(d - 1))
        if r < 0.9:
            return Not(tree(d - 1))
        return [IsNull(0), Like(3, "%a%"), Literal(bool(rng.random() < 0.5))][int(rng.integers(0, 3))]

    return [tree(3) for _ in range(k)]


@pytest.mark.parametrize("version", ["1.0", "2.6"])
def test_statistics_match_pyarrow(version):
    t = random_table(20000, 5)
    data = write(t, row_group_size=3000, version=version)
    f = ParquetFile(data)
    md = pq.ParquetFile(io.BytesIO(data)).metadata
    assert f.num_row_groups == md.num_row_groups and f.num_rows == md.num_rows
    assert [x[1] for x in f.fields] == [1, 2, 3, 4, 6, 5]
    for g in range(md.num_row_groups):
        assert f.row_group_rows(g) == md.row_group(g).num_rows
        for j in range(md.num_columns):
            ps = md.row_group(g).column(j).statistics
            s = f.statistics(g, j)
            assert (s is None) == (ps is None)
            if ps is None or not ps.has_min_max:
                continue
            lo, hi = ps.min, ps.max
            if isinstance(lo, str):
                lo, hi = lo.encode(), hi.encode()
            if hasattr(lo, "toordinal"):  # date32 statistics come back as dates
                epoch = 719163
                lo, hi = lo.toordinal() - epoch, hi.toordinal() - epoch
            assert (s.min, s.max) == (lo, hi)
            assert s.null_count == ps.null_count


def test_skipping_matches_oracle_and_is_sound():
    t = random_table(30000, 9)
    data = write(t, row_group_size=2500)
    f = ParquetFile(data)
    pf = pq.ParquetFile(io.BytesIO(data))
    names = t.schema.names
    rng = np.random.default_rng(17)
    skipped = 0
    for p in random_predicates(rng, 300):
        for g in range(f.num_row_groups):
            got = row_group_might_match(p, f, g)
            assert got == po.row_group_might_match(p, names, po.pyarrow_stats(pf.metadata, g)), p
            if not got:
                skipped += 1
                vals, valid = fo.eval_pred(pf.read_row_group(g).to_batches()[0], p)
                assert not np.any(vals & valid), p
    assert skipped > 100  # the rule did skip: the soundness check saw real cases


def test_integer_statistics_compare_as_f64():
    """`*v as f64` and `min as f64` (parquet_scan.rs:443-444): 2^53 + 3 rounds to
    2^53 + 4, so Le against a row group whose minimum is 2^53 + 4 might match."""
    big = 2**53 + 4
    f = ParquetFile(write(pa.table({"x": pa.array([big, big + 8], pa.int64())})))
    assert f.statistics(0, 0).min == big
    assert row_group_might_match(Le(0, ScalarValue.Int64(2**53 + 3)), f, 0)
    assert not row_group_might_match(Le(0, ScalarValue.Int64(2**53 - 1)), f, 0)
    assert po.stats_might_match("Le", "Int64", 2**53 + 3, ("INT64", big, big + 8))


def test_mismatched_scalar_kinds_might_match():
    """Only Int64/INT64, Int32/INT32, Float64/DOUBLE and Utf8/BYTE_ARRAY pairs
    consult statistics (parquet_scan.rs:437-469)."""
    st = ColumnStats(1, 10, 20, 0)  # INT32 statistics
    assert not stats_might_match("gt", ScalarValue.Int32(30), st)
    assert stats_might_match("gt", ScalarValue.Int64(30), st)
    assert stats_might_match("gt", ScalarValue.Date32(30), st)
    assert stats_might_match("eq", ScalarValue.Utf8("zz"), ColumnStats(6, None, b"b", 0))


def test_unsupported_leaves_and_bad_files():
    lib = L.load()
    t = pa.table({"f32": pa.array([1.0, 2.0], pa.float32()), "ts": pa.array([1, 2], pa.timestamp("ms")),
                  "ok": pa.array([1, 2], pa.int64())})
    f = ParquetFile(write(t))
    assert [x[1] for x in f.fields] == [0, 0, 2]  # float32 and timestamps are not decoded
    with pytest.raises(L.PrismError, match="nested"):
        ParquetFile(write(pa.table({"l": pa.array([[1, 2], [3]])})))
    with pytest.raises(L.PrismError, match="PAR1"):
        ParquetFile(b"PAR1" + b"\0" * 16)
    good = reference_bytes()
    with pytest.raises(L.PrismError, match="footer length"):
        ParquetFile(good[:4] + good[-8:])
    with pytest.raises(L.PrismError):
        ParquetFile(good[:-40] + good[-8:])
    assert lib.prism_parquet_open(None, 0, None) != 0
    ParquetFile(good).close()  # a successful call clears the thread's last error for later tests
