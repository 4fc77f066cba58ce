"""Parquet column chunks decoded into HBM (prism_parquet_read_device) against
pyarrow's reader of the same files: every writer layout the decoder accepts
(codecs, data page v1 / v2, dictionary and PLAIN pages, dictionary fallback
mid-chunk, RLE booleans, nulls, many small pages) must give the same values
and nulls, and the scan (parquet_scan.rs:55-258) the reference's row counts."""
import io

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

from gpu_util import say
from test_parquet import reference_bytes, reference_table
from prism_executor import _lib as L
from prism_executor.filter_project import Eq, Lt, ScalarValue
from prism_executor.parquet_scan import (ParquetFile, ParquetScanConfig, ParquetTable, parquet_scan,
                                         parquet_scan_device)

pytestmark = pytest.mark.gpu


def table(n, seed, null_rate=0.1):
    rng = np.random.default_rng(seed)
    words = np.array(["".join(chr(97 + c) for c in rng.integers(0, 26, rng.integers(0, 9))) for _ in range(40)])
    wide = ["w%07d" % v for v in rng.integers(0, 10**7, n)]  # too many values for one dictionary page
    m = lambda: rng.random(n) < null_rate
    return pa.table({
        "i32": pa.array(rng.integers(-2**31, 2**31, n).astype(np.int32), mask=m()),
        "i64": pa.array(rng.integers(-2**62, 2**62, n)),
        "f64": pa.array(rng.standard_normal(n), mask=m()),
        "s_low": pa.array(words[rng.integers(0, 40, n)], mask=m()),
        "s_wide": pa.array(wide, pa.utf8(), mask=m()),
        "b": pa.array(rng.random(n) < 0.3, mask=m()),
        "d": pa.array(rng.integers(0, 20000, n).astype(np.int32), mask=m()).cast(pa.date32()),
        "k": pa.array(rng.integers(0, 3, n).astype(np.int64)),  # long RLE runs of indices
    })


def write(t, **opts):
    buf = io.BytesIO()
    pq.write_table(t, buf, **opts)
    return buf.getvalue()


def check_file(ctx, data, columns=None):
    f = ParquetFile(data)
    ref = pq.ParquetFile(io.BytesIO(data))
    cols = list(range(f.num_columns)) if columns is None else columns
    t = ParquetTable(ctx, f, list(range(f.num_row_groups)), cols)
    try:
        assert t.n_batches == f.num_row_groups
        for g in range(f.num_row_groups):
            got = t.to_arrow(g)
            exp = ref.read_row_group(g, columns=[f.fields[c][0] for c in cols])
            assert got.num_rows == exp.num_rows
            for j, name in enumerate(exp.column_names):
                e = exp.column(name).combine_chunks()
                a = got.column(j)
                assert a.type == e.type, name
                assert a.null_count == e.null_count, name
                assert a.equals(e), (name, g)
    finally:
        t.close()


CASES = [
    dict(compression="snappy"),
    dict(compression="none"),
    dict(compression="gzip"),
    dict(compression="snappy", data_page_version="2.0"),
    dict(compression="none", data_page_version="2.0"),
    dict(compression="snappy", use_dictionary=False),
    dict(compression="snappy", use_dictionary=False, data_page_version="2.0"),
    dict(compression="snappy", data_page_size=2048),
    dict(compression="snappy", dictionary_pagesize_limit=4096),
]


@pytest.mark.parametrize("opts", CASES, ids=lambda o: ",".join(f"{k}={v}" for k, v in o.items()))
def test_layouts_decode_like_pyarrow(ctx, opts):
    say(f"parquet layout {opts}")
    check_file(ctx, write(table(30000, 1), row_group_size=12000, **opts))


def test_rle_booleans_and_required_columns(ctx):
    say("parquet rle booleans")
    t = table(20000, 2)
    t = t.append_column("b_req", pa.array(np.arange(20000) % 7 < 3))
    data = write(t, use_dictionary=["s_low"], column_encoding={"b": "RLE", "b_req": "RLE", "i64": "PLAIN"},
                 data_page_version="2.0", compression="snappy")
    check_file(ctx, data)


def test_projection_all_null_and_empty(ctx):
    say("parquet projection / all-null / empty")
    t = table(5000, 3)
    data = write(t, row_group_size=2000)
    check_file(ctx, data, columns=[6, 1, 3])
    nulls = pa.table({"x": pa.nulls(3000, pa.int64()), "s": pa.nulls(3000, pa.utf8()),
                      "y": pa.array(np.arange(3000), mask=np.ones(3000, bool))})
    check_file(ctx, write(nulls))
    check_file(ctx, write(table(0, 4)))
    one = pa.table({"s": pa.array(["only"] * 1000), "v": pa.array([5] * 1000, pa.int32())})  # bit width 0 indices
    check_file(ctx, write(one))


def test_large_chunk(ctx):
    say("parquet 2M rows")
    check_file(ctx, write(table(2_000_000, 5, null_rate=0.02), row_group_size=1_000_000, compression="snappy"))


def test_unsupported_column_fails_loudly(ctx):
    say("parquet unsupported leaf")
    f = ParquetFile(write(pa.table({"f32": pa.array([1.0, 2.0], pa.float32()), "ok": pa.array([1, 2])})))
    with pytest.raises(L.PrismError, match="f32"):
        ParquetTable(ctx, f, [0], [0, 1])
    check_file(ctx, f.data, columns=[1])


def test_reference_scan_expectations(ctx, tmp_path):
    """parquet_scan.rs tests: 200 rows for Lt(value, 150.0), 100 for
    Eq(name, "gamma"); the rows themselves equal the file's."""
    say("parquet reference scan")
    d = tmp_path / "ref"
    d.mkdir()
    (d / "test.parquet").write_bytes(reference_bytes())
    full = reference_table()
    got = parquet_scan(ParquetScanConfig([str(d)], predicate=Lt(1, ScalarValue.Float64(150.0))), ctx)
    assert sum(b.num_rows for b in got) == 200
    assert pa.Table.from_batches(got).equals(full.slice(0, 200).cast(got[0].schema))
    got = parquet_scan(ParquetScanConfig([str(d)], predicate=Eq(2, ScalarValue.Utf8("gamma"))), ctx)
    assert sum(b.num_rows for b in got) == 100
    assert got[0].column(2).to_pylist() == ["gamma"] * 100
    assert parquet_scan(ParquetScanConfig([]), ctx) == []
    # projection: non-projected columns come back as nulls (expand_projected_batches)
    got = parquet_scan(ParquetScanConfig([str(d)], projection=[0]), ctx)
    assert got[0].num_columns == 3 and got[0].column(1).null_count == got[0].num_rows
    assert pa.concat_arrays([b.column(0) for b in got]).to_pylist() == list(range(500))
    # target schema: reorder, missing column as nulls, int64 -> float64 cast
    target = pa.schema([pa.field("name", pa.utf8()), pa.field("extra", pa.int32()), pa.field("id", pa.float64())])
    got = parquet_scan(ParquetScanConfig([str(d)], target_schema=target), ctx)
    assert got[0].schema == target and got[0].column(1).null_count == got[0].num_rows
    assert got[0].column(2).to_pylist()[:3] == [0.0, 1.0, 2.0]
    tabs = parquet_scan_device(ParquetScanConfig([str(d)], predicate=Lt(0, ScalarValue.Int64(100))), ctx)
    assert len(tabs) == 1 and tabs[0].n_batches == 1
    n, cols = tabs[0].batch(0)
    assert n == 100 and cols[0].type == 2
    tabs[0].close()
