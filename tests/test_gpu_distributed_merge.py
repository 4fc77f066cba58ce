"""Reducer merge of partial aggregates on the GPU (prism_merge_partial_aggregates)
against the oracle's restatement and the reference's reducer test."""
import numpy as np
import pyarrow as pa
import pytest

from gpu_util import say
from test_distributed_merge import KATS, kat_partials

import oracle as orc
from prism_executor.distributed import expand_worker_aggregates, merge_partial_aggregates
from prism_executor.hash_aggregate import AggExpr, AggFunc

pytestmark = pytest.mark.gpu

FN = {"SUM": AggFunc.Sum, "COUNT": AggFunc.Count, "AVG": AggFunc.Avg, "MIN": AggFunc.Min, "MAX": AggFunc.Max,
      "COUNT_DISTINCT": AggFunc.CountDistinct}


def test_reference_reducer_test(ctx):
    say("distributed merge KAT")
    parts = kat_partials()
    orig = [AggExpr(1, AggFunc.Sum, "agg_0"), AggExpr(1, AggFunc.Count, "agg_1"), AggExpr(1, AggFunc.Avg, "agg_2")]
    out = merge_partial_aggregates(parts, 1, orig)
    assert out.schema.names == ["status", "agg_0", "agg_1", "agg_2"]
    assert [f.type for f in out.schema][1:] == [pa.float64(), pa.int64(), pa.float64()]
    rows = {s: (a, b, c) for s, a, b, c in zip(*[out.column(i).to_pylist() for i in range(4)])}
    assert out.column(0).to_pylist() == ["F", "O", "P"]
    assert rows["F"] == (130.0, 3, 130.0 / 3.0)
    assert rows["O"] == (50.0, 1, 50.0)
    assert rows["P"] == (70.0, 2, 35.0)


def partials(rng, n_parts, rows, original):
    """Producer outputs: group keys (Int64, Utf8), then the worker aggregates of
    `original` (SUM/MIN/MAX partials Float64 with a few nulls, COUNT partials
    Int64); one producer sends an empty batch."""
    worker, _ = orc.expand_worker_aggregates(original)
    out = []
    for p in range(n_parts):
        n = 0 if p == 1 else rows
        k1 = rng.integers(0, 40, n)
        k2 = np.array([b"ab", b"c", b"", b"xyz"], dtype=object)[rng.integers(0, 4, n)]
        cols = [pa.array(k1, pa.int64()), pa.array([s.decode() for s in k2], pa.utf8())]
        names = ["k1", "k2"]
        for i, (_, f) in enumerate(worker):
            if f in ("COUNT", "COUNT_DISTINCT"):
                cols.append(pa.array(rng.integers(0, 1000, n), pa.int64()))
            else:
                v = rng.integers(-10**6, 10**6, n).astype(np.float64) / 4.0  # exact in f64
                mask = rng.random(n) < 0.05
                cols.append(pa.array(v, pa.float64(), mask=mask))
            names.append(f"agg_{i}")
        out.append(pa.record_batch(cols, names=names))
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_random_partials_match_oracle(ctx, seed):
    say(f"distributed merge random seed={seed}")
    rng = np.random.default_rng(seed)
    original = [(2, "SUM"), (3, "COUNT"), (4, "AVG"), (5, "MIN"), (6, "MAX"), (7, "COUNT_DISTINCT"), (2, "AVG")]
    parts = partials(rng, 4, 3000, original)
    first, vals = orc.reducer_merge(parts, 2, original)
    orig = [AggExpr(c, FN[f], f"o{i}") for i, (c, f) in enumerate(original)]
    out = merge_partial_aggregates(parts, 2, orig)
    assert out.num_rows == len(first)
    exp_k1 = [parts[b].column(0)[r].as_py() for b, r in first]
    exp_k2 = [parts[b].column(1)[r].as_py() for b, r in first]
    assert out.column(0).to_pylist() == exp_k1 and out.column(1).to_pylist() == exp_k2
    for j, (_, f) in enumerate(original):
        got = out.column(2 + j).to_pylist()
        if f in ("COUNT", "COUNT_DISTINCT"):
            assert out.schema.field(2 + j).type == pa.int64()
            assert got == vals[j]
        else:
            assert out.schema.field(2 + j).type == pa.float64()
            np.testing.assert_allclose(got, vals[j], rtol=1e-12, atol=0)


def test_no_partials_rows(ctx):
    say("distributed merge empty")
    rng = np.random.default_rng(3)
    original = [(2, "AVG"), (4, "COUNT")]
    parts = partials(rng, 2, 0, original)
    orig = [AggExpr(c, FN[f], f"o{i}") for i, (c, f) in enumerate(original)]
    out = merge_partial_aggregates(parts, 2, orig)
    assert out.num_rows == 0 and out.schema.names == ["k1", "k2", "o0", "o1"]
