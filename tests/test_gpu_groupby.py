"""Hash aggregation on the GPU against the oracle (hash_aggregate.rs:182-447):
groups in first-appearance order, FxHash group ids, exact integer sums.  Every
test runs twice: with the update kernels generated per plan (hipRTC) and with
the prebuilt interpreting kernels (PRISM_JIT=0)."""
import math

import numpy as np
import pyarrow as pa
import pytest

from gpu_util import exact_sum, say, to_dev

import oracle as orc
from prism_executor import _lib as L
from prism_executor import datagen as dg
from prism_executor import device as D

pytestmark = pytest.mark.gpu

I64, I32, F64, UTF8 = L.TYPE_INT64, L.TYPE_INT32, L.TYPE_FLOAT64, L.TYPE_UTF8


@pytest.fixture(autouse=True, params=["generated", "prebuilt"])
def engine(request, monkeypatch):
    monkeypatch.setenv("PRISM_JIT", "1" if request.param == "generated" else "0")
    say("engine", request.param)
    return request.param


def gpu_groupby(ctx, types, cols, keys, values, aggs, n, filt=None, strings=(), batches=None, expected=0):
    g = D.GroupBy(ctx, types, keys, values, aggs, filter_prog=filt, strings=strings, expected_groups=expected)
    if batches is None:
        g.update(cols, n, 0)
    else:
        for c, m, base in batches:
            g.update(c, m, base)
    k = g.finish()
    out = g.read(k)
    g.close()
    return out


def test_config0_groupby_bit_exact(ctx):
    say("config0 1M rows")
    d = dg.groupby_small()
    kd, vd = to_dev(d["key"]), to_dev(d["value"])
    out = gpu_groupby(ctx, [I64, I64], [D.col_i64(kd), D.col_i64(vd)], [0], [[("LOAD_I", 1)]],
                      [("SUM", 0), ("COUNT", 0), ("MIN", 0), ("MAX", 0)], len(d["key"]))
    rb = pa.record_batch([pa.array(d["key"]), pa.array(d["value"])], names=["k", "v"])
    ref = orc.hash_aggregate([rb], [0], [(1, "SUM"), (1, "COUNT"), (1, "MIN"), (1, "MAX")])
    assert out["first_row"].tolist() == ref["first_row"].tolist()
    assert out["key_hash"].tolist() == ref["key_hash"].tolist()
    for a in (0, 2, 3):
        assert np.array_equal(out["f64"][a], ref["f64"][a])
    assert np.array_equal(out["i64"][1], ref["i64"][1])
    assert out["sum128"][0] == ref["sum128"][0]


def test_multibatch_row_base_first_appearance(ctx):
    say("multi-batch")
    d = dg.groupby_small(300_000)
    cuts = [0, 1, 77_777, 150_000, 150_001, 300_000]
    batches, rbs = [], []
    for a, b in zip(cuts[:-1], cuts[1:]):
        kd, vd = to_dev(d["key"][a:b]), to_dev(d["value"][a:b])
        batches.append(([D.col_i64(kd), D.col_i64(vd)], b - a, a))
        rbs.append(pa.record_batch([pa.array(d["key"][a:b]), pa.array(d["value"][a:b])], names=["k", "v"]))
        batches[-1][0].append((kd, vd))  # keep tensors alive
    bt = [(c[:2], m, base) for c, m, base in batches]
    out = gpu_groupby(ctx, [I64, I64], None, [0], [[("LOAD_I", 1)]], [("SUM", 0), ("MAX", 0)], 0, batches=bt)
    ref = orc.hash_aggregate(rbs, [0], [(1, "SUM"), (1, "MAX")], n_threads=2)
    glob = [cuts[fb] + fr for fb, fr in zip(ref["first_batch"], ref["first_row"])]
    assert out["first_row"].tolist() == glob
    assert np.array_equal(out["f64"][0], ref["f64"][0]) and np.array_equal(out["f64"][1], ref["f64"][1])


def test_utf8_int32_keys_nulls_and_f64_values(ctx):
    say("utf8/int32 keys, nulls, f64 values")
    rng = np.random.default_rng(1)
    n = 200_000
    words = ["", "a", "a\x00", "abcdefgh", "abcdefghi", "zz", "été"]
    s = pa.array([words[i] if i % 11 else None for i in rng.integers(0, len(words), n)])
    k32 = pa.array(rng.integers(-5, 5, n).astype(np.int32))
    vals = rng.standard_normal(n) * 1e3
    vals[::1000] = np.nan
    v = pa.array(vals)
    rb = pa.record_batch([s, k32, v], names=["s", "k", "v"])
    sc = orc.from_arrow(s)
    valid = np.packbits(np.asarray(s.is_valid()), bitorder="little")
    data_d, off_d, val_d = to_dev(sc.data if len(sc.data) else np.zeros(1, np.uint8)), to_dev(sc.offsets), to_dev(valid)
    kd, vd = to_dev(np.asarray(k32)), to_dev(vals)
    cols = [D.col_utf8(data_d, off_d, val_d), D.col_i32(kd), D.col_f64(vd)]
    out = gpu_groupby(ctx, [UTF8, I32, F64], cols, [0, 1], [[("LOAD_F", 2)]],
                      [("SUM", 0), ("COUNT", 0), ("MIN", 0), ("MAX", 0), ("AVG", 0)], n)
    ref = orc.hash_aggregate([rb], [0, 1], [(2, "SUM"), (2, "COUNT"), (2, "MIN"), (2, "MAX"), (2, "AVG")])
    assert out["first_row"].tolist() == ref["first_row"].tolist()
    assert out["key_hash"].tolist() == ref["key_hash"].tolist()
    assert np.array_equal(out["i64"][1], ref["i64"][1])
    assert np.array_equal(out["f64"][2], ref["f64"][2]) and np.array_equal(out["f64"][3], ref["f64"][3])
    # f64 sums: order-dependent; compare with the exactly rounded sum per group
    ob = orc.Batch.from_arrow(rb)
    kh = np.array([orc.group_hash(ob, [0, 1], r) for r in range(n)], np.uint64)
    for i, h in enumerate(ref["key_hash"]):
        sel = vals[kh == h]
        if np.isnan(sel).any():
            assert math.isnan(out["f64"][0][i])
            continue
        exact = math.fsum(sel)
        assert abs(out["f64"][0][i] - exact) <= 1e-12 * max(math.fsum(abs(sel)), 1e-300)


def test_global_aggregates_skip_nulls(ctx):
    say("global aggregates")
    v = pa.array([None, 5, None, 2, 7, None], pa.int64())
    rb = pa.record_batch([v], names=["v"])
    c = orc.from_arrow(v)
    valid = np.packbits(np.asarray(v.is_valid()), bitorder="little")
    vd, md = to_dev(c.data), to_dev(valid)
    out = gpu_groupby(ctx, [I64], [D.col_i64(vd, md)], [], [[("LOAD_I", 0)]],
                      [("COUNT", 0), ("SUM", 0), ("AVG", 0), ("MIN", 0), ("MAX", 0)], len(v))
    ref = orc.hash_aggregate([rb], [], [(0, "COUNT"), (0, "SUM"), (0, "AVG"), (0, "MIN"), (0, "MAX")])
    assert out["i64"][0][0] == ref["i64"][0][0] == 3
    for a in (1, 2, 3, 4):
        assert out["f64"][a][0] == ref["f64"][a][0]


def _config2_cols(n, seed=dg.SEED):
    d = dg.filter_groupby(n, seed)
    dev = {k: (to_dev(v[0]), to_dev(v[1])) if isinstance(v, tuple) else to_dev(v) for k, v in d.items()}
    cols = [D.col_i32(dev["date"]), D.col_i64(dev["price"]), D.col_i64(dev["discount"]), D.col_i32(dev["quantity"]),
            D.col_utf8(*dev["flag_a"]), D.col_utf8(*dev["flag_b"])]
    return d, dev, cols


TYPES2 = [I32, I64, I64, I32, UTF8, UTF8]
FILTER2 = [("LOAD_I", 0), ("CONST_I", 0, 2470), ("CMP_I", L.CMP["LE"])]
EXACT2 = [("LOAD_I", 1), ("CONST_I", 0, 100), ("LOAD_I", 2), ("SUB_I",), ("MUL_I",)]
F64_2 = [("LOAD_I", 1), ("I2F",), ("CONST_F", 0, 0, 100.0), ("LOAD_I", 2), ("I2F",), ("SUB_F",), ("MUL_F",)]
FRAC2 = [("LOAD_I", 1), ("I2F",), ("CONST_F", 0, 0, 0.01), ("MUL_F",), ("CONST_F", 0, 0, 1.0), ("LOAD_I", 2),
         ("I2F",), ("CONST_F", 0, 0, 0.01), ("MUL_F",), ("SUB_F",), ("MUL_F",)]


def test_fused_filter_groupby_exact_int128(ctx):
    say("fused filter+groupby exact")
    n = 2_000_000
    d, dev, cols = _config2_cols(n)
    out = gpu_groupby(ctx, TYPES2, cols, [4, 5], [EXACT2, F64_2],
                      [("SUM", 0), ("COUNT", 0), ("AVG", 0), ("SUM", 1)], n, filt=FILTER2)
    keep = d["date"] <= 2470
    rev = d["price"] * (100 - d["discount"])
    a, b = d["flag_a"][0], d["flag_b"][0]
    assert out["row_count"].sum() == keep.sum()
    order = []
    for i, fr in enumerate(out["first_row"]):
        ga, gb = a[fr], b[fr]
        sel = keep & (a == ga) & (b == gb)
        order.append(np.nonzero(sel)[0][0])
        s = exact_sum(rev[sel])
        assert out["sum128"][0][i] == s
        assert out["i64"][1][i] == sel.sum()
        assert out["f64"][0][i] == float(s)  # correctly rounded exact sum
        assert out["f64"][2][i] == float(s) / float(sel.sum())
        assert out["f64"][3][i] == float(s)  # integral f64 values are summed exactly
    assert order == sorted(order) and len(order) == 6


def test_fused_filter_groupby_float_variant(ctx):
    say("fused filter+groupby f64")
    n = 1_000_000
    d, dev, cols = _config2_cols(n, seed=77)
    out = gpu_groupby(ctx, TYPES2, cols, [4, 5], [FRAC2], [("SUM", 0), ("AVG", 0)], n, filt=FILTER2)
    keep = d["date"] <= 2470
    rev = (d["price"] * 0.01) * (1.0 - d["discount"] * 0.01)
    a, b = d["flag_a"][0], d["flag_b"][0]
    for i, fr in enumerate(out["first_row"]):
        sel = keep & (a == a[fr]) & (b == b[fr])
        exact = math.fsum(rev[sel])
        assert abs(out["f64"][0][i] - exact) <= 1e-12 * abs(exact)


def test_high_cardinality_growth(ctx):
    say("high cardinality growth")
    n = 1_000_000
    key = dg.uniform_i64(n, 5, 1, 0, 400_000)
    val = dg.uniform_i64(n, 5, 2, -1000, 2000)
    kd, vd = to_dev(key), to_dev(val)
    out = gpu_groupby(ctx, [I64, I64], [D.col_i64(kd), D.col_i64(vd)], [0], [[("LOAD_I", 1)]],
                      [("SUM", 0), ("COUNT", 0)], n)
    rb = pa.record_batch([pa.array(key), pa.array(val)], names=["k", "v"])
    ref = orc.hash_aggregate([rb], [0], [(1, "SUM"), (1, "COUNT")])
    assert out["first_row"].tolist() == ref["first_row"].tolist()
    assert np.array_equal(out["f64"][0], ref["f64"][0]) and np.array_equal(out["i64"][1], ref["i64"][1])


def test_reset_and_reuse(ctx):
    say("reset/reuse")
    d = dg.groupby_small(100_000)
    kd, vd = to_dev(d["key"]), to_dev(d["value"])
    g = D.GroupBy(ctx, [I64, I64], [0], [[("LOAD_I", 1)]], [("SUM", 0)])
    res = []
    for _ in range(2):
        g.update([D.col_i64(kd), D.col_i64(vd)], 100_000, 0)
        k = g.finish()
        res.append(g.read(k))
        g.reset()
    g.close()
    assert res[0]["sum128"][0] == res[1]["sum128"][0]


def test_exact_overflow_is_an_error(ctx):
    say("exact overflow error")
    x = to_dev(np.array([2**62, 3], np.int64))
    with pytest.raises(L.PrismError) as e:
        gpu_groupby(ctx, [I64], [D.col_i64(x)], [], [[("LOAD_I", 0), ("CONST_I", 0, 4), ("MUL_I",)]],
                    [("SUM", 0)], 2)
    assert e.value.kind == "Execution"


# The forms below pick the thread-private LDS accumulators (a small expected
# group count and few accumulator words), its overflow into the global table
# (expected count below the real one), and the prefix that measures the group
# count when none is given (more than two sample prefixes of rows).
@pytest.mark.parametrize("n,expected", [(2_000_000, 8), (2_000_000, 4), (3_000_000, 0), (2_000_000, 2)])
def test_config2_shape_all_forms(ctx, n, expected):
    say(f"config2 shape n={n} expected={expected}")
    d, dev, cols = _config2_cols(n, seed=5 + expected)
    out = gpu_groupby(ctx, TYPES2, cols, [4, 5], [EXACT2], [("SUM", 0), ("COUNT", 0), ("AVG", 0)], n,
                      filt=FILTER2, expected=expected)
    keep = d["date"] <= 2470
    rev = d["price"] * (100 - d["discount"])
    a, b = d["flag_a"][0], d["flag_b"][0]
    assert out["row_count"].sum() == keep.sum() and len(out["first_row"]) == 6
    firsts = []
    for i, fr in enumerate(out["first_row"]):
        sel = keep & (a == a[fr]) & (b == b[fr])
        firsts.append(int(np.nonzero(sel)[0][0]))
        s = exact_sum(rev[sel])
        assert out["sum128"][0][i] == s
        assert out["i64"][1][i] == sel.sum()
        assert out["f64"][2][i] == float(s) / float(sel.sum())
    assert firsts == list(out["first_row"]) and firsts == sorted(firsts)


def test_private_form_nullable_keys(ctx):
    say("private form, nullable utf8 + int32 keys")
    rng = np.random.default_rng(11)
    n = 300_000
    words = ["yy", None]
    s = pa.array([words[i] for i in rng.integers(0, 2, n)])
    k32 = pa.array([None if i == 2 else int(i) for i in rng.integers(0, 3, n)], pa.int32())
    val = rng.integers(-(1 << 40), 1 << 40, n)
    rb = pa.record_batch([s, k32, pa.array(val)], names=["s", "k", "v"])
    sc = orc.from_arrow(s)
    sv = np.packbits(np.asarray(s.is_valid()), bitorder="little")
    kv = np.packbits(np.asarray(k32.is_valid()), bitorder="little")
    kvals = np.asarray(k32.fill_null(0)).astype(np.int32)
    data_d, off_d, sv_d = to_dev(sc.data), to_dev(sc.offsets), to_dev(sv)
    kd, kv_d, vd = to_dev(kvals), to_dev(kv), to_dev(val)
    cols = [D.col_utf8(data_d, off_d, sv_d), D.col_i32(kd, kv_d), D.col_i64(vd)]
    out = gpu_groupby(ctx, [UTF8, I32, I64], cols, [0, 1], [[("LOAD_I", 2)]], [("SUM", 0), ("COUNT", 0)], n,
                      expected=6)
    ref = orc.hash_aggregate([rb], [0, 1], [(2, "SUM"), (2, "COUNT")])
    assert out["first_row"].tolist() == ref["first_row"].tolist()
    assert out["key_hash"].tolist() == ref["key_hash"].tolist()
    assert np.array_equal(out["i64"][1], ref["i64"][1])
    assert list(out["sum128"][0]) == list(ref["sum128"][0])


def test_private_form_global_aggregates(ctx):
    say("private form, global aggregates with nulls")
    rng = np.random.default_rng(12)
    n = 500_000
    vals = rng.integers(-10**12, 10**12, n)
    valid = rng.random(n) > 0.1
    v = pa.array(vals, mask=~valid)
    rb = pa.record_batch([v], names=["v"])
    vd, md = to_dev(vals), to_dev(np.packbits(valid, bitorder="little"))
    out = gpu_groupby(ctx, [I64], [D.col_i64(vd, md)], [], [[("LOAD_I", 0)]],
                      [("COUNT", 0), ("SUM", 0), ("AVG", 0), ("MIN", 0), ("MAX", 0)], n, expected=1)
    ref = orc.hash_aggregate([rb], [], [(0, "COUNT"), (0, "SUM"), (0, "AVG"), (0, "MIN"), (0, "MAX")])
    assert out["i64"][0][0] == ref["i64"][0][0] == int(valid.sum())
    assert out["sum128"][1][0] == exact_sum(vals[valid])
    for a in (3, 4):
        assert out["f64"][a][0] == ref["f64"][a][0]


def test_generated_kernels_ran_without_compile_failures(ctx, engine):
    st = D.jit_stats()
    say("jit stats", st)
    assert st["failures"] == 0, st["last_error"]
    assert st["kernels"] > 0
