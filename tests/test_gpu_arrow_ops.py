"""The Arrow-level operator API on the GPU, written like the reference's own
tests (hash_join.rs:672-744, hash_aggregate.rs:529-661) and checked against the
oracle for the cases those tests leave open."""
import json
import os

import numpy as np
import pyarrow as pa
import pytest

from gpu_util import say

import oracle as orc
from prism_executor._lib import PrismError
from prism_executor.hash_aggregate import AggExpr, AggFunc, HashAggConfig, hash_aggregate, hash_aggregate_batches
from prism_executor.hash_join import (HashJoinConfig, HashTable, JoinType, hash_join, hash_join_probe_chunked,
                                      probe_matches_chunked)

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def rb(d, types=None):
    types = types or {}
    return pa.record_batch([pa.array(v, types.get(k)) for k, v in d.items()], names=list(d.keys()))


def build_side():
    return rb(KATS["hash_join"]["build"])


def probe_side():
    return rb(KATS["hash_join"]["probe"])


def test_inner_join():
    say("arrow inner join")
    cfg = HashJoinConfig(JoinType.Inner, build_keys=[0], probe_keys=[2])
    r = hash_join(probe_side(), build_side(), cfg)
    assert r.num_rows == 3 and r.num_columns == 5
    assert r.schema.names == ["emp_id", "name", "dept_id", "dept_id_right", "dept_name"]
    assert r.column(0).to_pylist() == [100, 101, 102]
    assert r.column(4).to_pylist() == ["Engineering", "Sales", "Engineering"]


def test_left_semi_and_anti_join():
    say("arrow semi/anti")
    r = hash_join(probe_side(), build_side(), HashJoinConfig(JoinType.LeftSemi, [0], [2]))
    assert r.num_rows == 3 and r.num_columns == 3
    r = hash_join(probe_side(), build_side(), HashJoinConfig(JoinType.LeftAnti, [0], [2]))
    assert r.num_rows == 1 and r.column(1).to_pylist() == ["Dave"]


def test_left_join_pads_with_build_row_zero():
    say("arrow left join quirk")
    r = hash_join(probe_side(), build_side(), HashJoinConfig(JoinType.Left, [0], [2]))
    # matched pairs first, then unmatched probe rows paired with build row 0 (hash_join.rs:472-477)
    assert r.column(0).to_pylist() == [100, 101, 102, 103]
    assert r.column(4).to_pylist() == ["Engineering", "Sales", "Engineering", "Engineering"]
    assert r.column(4).null_count == 0


def test_right_and_full_join():
    say("arrow right/full")
    r = hash_join(probe_side(), build_side(), HashJoinConfig(JoinType.Right, [0], [2]))
    assert r.num_rows == 4
    assert r.column(0).to_pylist() == [100, 101, 102, None]
    assert r.column(4).to_pylist() == ["Engineering", "Sales", "Engineering", "Marketing"]
    assert r.schema.field(0).nullable
    f = hash_join(probe_side(), build_side(), HashJoinConfig(JoinType.Full, [0], [2]))
    assert f.num_rows == 5
    assert f.column(0).to_pylist() == [100, 101, 102, 103, None]
    assert f.column(3).to_pylist() == [1, 2, 1, None, 3]


def test_general_path_multikey_with_nulls_matches_oracle():
    say("arrow general path")
    rng = np.random.default_rng(2)
    n_b, n_p = 3000, 20000
    bs = pa.array([None if i % 17 == 0 else f"k{i}" for i in rng.integers(0, 40, n_b)])
    bi = pa.array(rng.integers(0, 5, n_b).astype(np.int32))
    ps = pa.array([None if i % 13 == 0 else f"k{i}" for i in rng.integers(0, 50, n_p)])
    pi_ = pa.array(rng.integers(0, 6, n_p).astype(np.int32))
    build = pa.record_batch([bs, bi, pa.array(np.arange(n_b))], names=["s", "i", "row"])
    probe = pa.record_batch([ps, pi_, pa.array(np.arange(n_p))], names=["s", "i", "prow"])
    t = HashTable.build(build, [0, 1])
    gp, gb = t.probe(probe, [0, 1])
    op, ob = orc.join_probe(build, [0, 1], probe, [0, 1])
    assert gp.to_numpy().tolist() == op.tolist() and gb.to_numpy().tolist() == ob.tolist()


def test_int64_fast_path_matches_oracle_with_nulls():
    say("arrow int64 fast path nulls")
    b = pa.record_batch([pa.array([0, 5, 7, 5], pa.int64())], names=["k"])
    p = pa.record_batch([pa.array([None, 5, 9, 7, 0], pa.int64())], names=["k"])
    t = HashTable.build(b, [0])
    gp, gb = t.probe(p, [0])
    op, ob = orc.join_probe(b, [0], p, [0])
    assert gp.to_pylist() == op.tolist() and gb.to_pylist() == ob.tolist()


def test_chunked_probe_and_join():
    say("arrow chunked")
    rng = np.random.default_rng(4)
    build = pa.record_batch([pa.array(np.arange(0, 2000, 2, dtype=np.int64)), pa.array(np.arange(1000))],
                            names=["k", "v"])
    probes = [pa.record_batch([pa.array(rng.integers(0, 3000, n).astype(np.int64))], names=["k"])
              for n in (1000, 0, 1, 2500)]
    ms = probe_matches_chunked(probes, build, [0], [0])
    assert len(ms) == 4
    for m, p in zip(ms, probes):
        op, ob = orc.join_probe(build, [0], p, [0])
        assert m.probe_indices.to_pylist() == op.tolist() and m.build_indices.to_pylist() == ob.tolist()
    for jt in JoinType:
        outs = hash_join_probe_chunked(probes, build, HashJoinConfig(jt, [0], [0]))
        ref = orc.hash_join_probe_chunked(probes, build, int(jt), [0], [0])
        assert len(outs) == len(ref), jt
        for o, r in zip(outs, ref):
            assert o.schema.equals(r.schema), jt
            assert o.equals(r), jt


def test_key_type_mismatch_is_a_join_error():
    say("arrow key mismatch")
    b = pa.record_batch([pa.array([1, 2], pa.int64())], names=["k"])
    p = pa.record_batch([pa.array([1, 2], pa.int32())], names=["k"])
    with pytest.raises(PrismError) as e:
        hash_join(p, b, HashJoinConfig(JoinType.Inner, [0], [0]))
    assert e.value.kind == "Join"


def sales_batch():
    return rb(KATS["hash_aggregate"]["sales"])


def test_sum_by_region():
    say("arrow agg sum by region")
    cfg = HashAggConfig(group_by=[0], aggregates=[AggExpr(2, AggFunc.Sum, "total_amount")])
    r = hash_aggregate(sales_batch(), cfg)
    assert r.num_rows == 2 and r.num_columns == 2
    assert sum(r.column(1).to_pylist()) == 800.0
    assert r.schema.field(1).type == pa.float64() and r.schema.names == ["region", "total_amount"]
    assert r.column(0).to_pylist() == ["east", "west"]


def test_count_and_avg():
    say("arrow agg count/avg")
    cfg = HashAggConfig([0], [AggExpr(2, AggFunc.Count, "cnt"), AggExpr(2, AggFunc.Avg, "avg_amount")])
    r = hash_aggregate(sales_batch(), cfg)
    assert r.num_rows == 2 and r.num_columns == 3
    assert r.schema.field(1).type == pa.int64()
    assert r.column(1).to_pylist() == [3, 2] and r.column(2).to_pylist() == [100.0, 250.0]


def test_global_count_distinct_preserves_large_int64_values():
    say("arrow agg global distinct")
    b = pa.record_batch([pa.array([2**53, 2**53 + 1, 2**53, None], pa.int64())], names=["id"])
    r = hash_aggregate(b, HashAggConfig([], [AggExpr(0, AggFunc.CountDistinct, "distinct_ids")]))
    assert r.column(0).to_pylist() == [2]


def test_grouped_count_distinct_supports_utf8():
    say("arrow agg grouped distinct")
    g = KATS["hash_aggregate"]["grouped_count_distinct_utf8"]
    b = rb({"region": g["region"], "label": g["label"]})
    r = hash_aggregate(b, HashAggConfig([0], [AggExpr(1, AggFunc.CountDistinct, "distinct_labels")]))
    assert dict(zip(r.column(0).to_pylist(), r.column(1).to_pylist())) == {"east": 2, "west": 2}


def test_batches_match_oracle_and_keys_come_from_first_rows():
    say("arrow agg batches vs oracle")
    rng = np.random.default_rng(9)
    batches = []
    for n in (5000, 0, 3000, 7000):
        keys = pa.array([None if x == 0 else f"g{x}" for x in rng.integers(0, 30, n)], pa.string())
        k2 = pa.array(rng.integers(-3, 3, n).astype(np.int32))
        v = pa.array(rng.integers(-10**12, 10**12, n))
        f = pa.array(rng.standard_normal(n))
        batches.append(pa.record_batch([keys, k2, v, f], names=["k", "k2", "v", "f"]))
    cfg = HashAggConfig([0, 1], [AggExpr(2, AggFunc.Sum, "s"), AggExpr(2, AggFunc.Min, "mn"),
                                 AggExpr(2, AggFunc.Max, "mx"), AggExpr(3, AggFunc.Count, "c"),
                                 AggExpr(2, AggFunc.CountDistinct, "d")])
    r = hash_aggregate_batches(batches, cfg)
    ref = orc.hash_aggregate(batches, [0, 1], [(2, "SUM"), (2, "MIN"), (2, "MAX"), (3, "COUNT"),
                                               (2, "COUNT_DISTINCT")])
    assert r.num_rows == ref["n_groups"]
    exp_keys = [batches[b].column(0)[i].as_py() for b, i in zip(ref["first_batch"], ref["first_row"])]
    assert r.column(0).to_pylist() == exp_keys
    assert r.column(3).to_pylist() == ref["f64"][1].tolist()
    assert r.column(4).to_pylist() == ref["f64"][2].tolist()
    assert r.column(5).to_pylist() == ref["i64"][3].tolist()
    assert r.column(6).to_pylist() == ref["i64"][4].tolist()
    # SUM: the exactly rounded value of the exact sum
    assert r.column(2).to_pylist() == [float(s) for s in ref["sum128"][0]]


def test_exact_decimal_sums():
    say("arrow agg exact decimal")
    vals = [2**62, 2**62, 2**62, -5]
    b = pa.record_batch([pa.array([1, 1, 1, 2]), pa.array(vals, pa.int64())], names=["k", "v"])
    r = hash_aggregate(b, HashAggConfig([0], [AggExpr(1, AggFunc.Sum, "s")]), exact=True)
    assert r.schema.field(1).type == pa.decimal128(38, 0)
    assert [int(x) for x in r.column(1).to_pylist()] == [3 * 2**62, -5]


def test_empty_input_gives_empty_result():
    say("arrow agg empty")
    b = pa.record_batch([pa.array([], pa.int64()), pa.array([], pa.float64())], names=["k", "v"])
    r = hash_aggregate(b, HashAggConfig([0], [AggExpr(1, AggFunc.Sum, "s"), AggExpr(1, AggFunc.Count, "c")]))
    assert r.num_rows == 0 and r.schema.names == ["k", "s", "c"]
    assert r.schema.field(2).type == pa.int64()


def random_sides(seed, n_b, n_p, domain_b, domain_p, nulls=True):
    rng = np.random.default_rng(seed)
    bk = pa.array(rng.integers(0, domain_b, n_b), mask=(rng.random(n_b) < 0.05) if nulls else None)
    pk = pa.array(rng.integers(0, domain_p, n_p), mask=(rng.random(n_p) < 0.05) if nulls else None)
    bs = pa.array([None if i % 11 == 0 else f"b{i}" for i in rng.integers(0, 100, n_b)], pa.string())
    ps = pa.array([None if i % 13 == 0 else f"p{i}" for i in rng.integers(0, 100, n_p)], pa.string())
    build = pa.record_batch([bk, bs, pa.array(rng.standard_normal(n_b))], names=["k", "s", "x"])
    probe = pa.record_batch([pk, ps, pa.array(rng.integers(-5, 5, n_p).astype(np.int32))], names=["k", "s", "y"])
    return build, probe


@pytest.mark.parametrize("jt", list(JoinType))
def test_every_join_type_matches_oracle_on_random_inputs(jt):
    say("arrow join type vs oracle", jt.name)
    for seed, (nb, npr, db, dp) in enumerate([(3000, 9000, 1500, 2500), (1, 500, 3, 3), (2000, 0, 10, 10)]):
        build, probe = random_sides(seed, nb, npr, db, dp)
        got = hash_join(probe, build, HashJoinConfig(jt, [0], [0]))
        ref = orc.hash_join(probe, build, int(jt), [0], [0])
        assert got.schema.equals(ref.schema), (jt, seed)
        assert got.equals(ref), (jt, seed)
    # general (multi-key, Utf8 + Int64) path
    build, probe = random_sides(7, 2500, 6000, 40, 50)
    got = hash_join(probe, build, HashJoinConfig(jt, [1, 0], [1, 0]))
    ref = orc.hash_join(probe, build, int(jt), [1, 0], [1, 0])
    assert got.equals(ref) and got.schema.equals(ref.schema)


def test_left_join_with_empty_build_is_refused():
    say("arrow left join, empty build")
    build = pa.record_batch([pa.array([], pa.int64()), pa.array([], pa.string())], names=["k", "s"])
    probe = pa.record_batch([pa.array([1, 2], pa.int64())], names=["k"])
    with pytest.raises(PrismError) as e:
        hash_join(probe, build, HashJoinConfig(JoinType.Left, [0], [0]))
    assert e.value.kind == "Arrow"
    with pytest.raises(PrismError):
        hash_join_probe_chunked([probe], build, HashJoinConfig(JoinType.Left, [0], [0]))
    # nothing to pad: an empty probe side is fine, and so are the other types
    empty_probe = probe.slice(0, 0)
    assert hash_join(empty_probe, build, HashJoinConfig(JoinType.Left, [0], [0])).num_rows == 0
    assert hash_join(probe, build, HashJoinConfig(JoinType.Full, [0], [0])).num_rows == 2


def test_count_distinct_float_zero_and_nan_folding():
    say("arrow distinct float +-0 / NaN")
    rng = np.random.default_rng(12)
    n = 20_000
    nan2 = np.frombuffer(np.uint64(0x7ff8dead00000001).tobytes(), np.float64)[0]
    pool = np.array([0.0, -0.0, np.nan, nan2, 1.5, -1.5, 2.0**60, 5e-324])
    vals = pool[rng.integers(0, len(pool), n)]
    b = pa.record_batch([pa.array(rng.integers(0, 50, n)), pa.array(vals, mask=rng.random(n) < 0.1)],
                        names=["k", "f"])
    r = hash_aggregate(b, HashAggConfig([0], [AggExpr(1, AggFunc.CountDistinct, "d")]))
    ref = orc.hash_aggregate([b], [0], [(1, "COUNT_DISTINCT")])
    assert r.column(1).to_pylist() == ref["i64"][0].tolist()
    g = hash_aggregate(b, HashAggConfig([], [AggExpr(1, AggFunc.CountDistinct, "d")]))
    assert g.column(0).to_pylist() == [6]  # {0, NaN, 1.5, -1.5, 2^60, 5e-324}


@pytest.mark.parametrize("fp_bits", [64, 3, 0])
def test_utf8_count_distinct_is_exact_under_fingerprint_collisions(fp_bits, monkeypatch):
    say("arrow distinct utf8, fingerprint bits", fp_bits)
    monkeypatch.setenv("PRISM_DISTINCT_FP_BITS", str(fp_bits))
    rng = np.random.default_rng(13)
    words = [f"w{i}" for i in range(300)] + ["", "a\x00", "a"]
    batches = []
    for n in (4000, 0, 2500, 3000):
        ws = [words[i] for i in rng.integers(0, len(words), n)]
        batches.append(pa.record_batch([pa.array(rng.integers(0, 9, n).astype(np.int32)),
                                        pa.array(ws, pa.string(), mask=rng.random(n) < 0.05)], names=["k", "s"]))
    r = hash_aggregate_batches(batches, HashAggConfig([0], [AggExpr(1, AggFunc.CountDistinct, "d")]))
    ref = orc.hash_aggregate(batches, [0], [(1, "COUNT_DISTINCT")])
    assert r.column(1).to_pylist() == ref["i64"][0].tolist()


def test_concurrent_callers_on_one_context():
    """Four threads run joins and aggregations through one process-wide
    context at once (the reference's executor calls operators from rayon
    par_iter); every result equals the oracle's."""
    import threading
    say("arrow concurrent callers")
    errors = []

    def worker(t):
        try:
            for it in range(6):
                seed = 100 * t + it
                build, probe = random_sides(seed, 4000, 30_000, 3000, 5000, nulls=False)
                got = hash_join(probe, build, HashJoinConfig(JoinType.Inner, [0], [0]))
                assert got.equals(orc.hash_join(probe, build, 0, [0], [0])), ("join", seed)
                rng = np.random.default_rng(seed)
                n = 50_000
                b = pa.record_batch([pa.array(rng.integers(0, 500, n)), pa.array(rng.integers(-10**9, 10**9, n))],
                                    names=["k", "v"])
                r = hash_aggregate(b, HashAggConfig([0], [AggExpr(1, AggFunc.Sum, "s"),
                                                          AggExpr(1, AggFunc.Count, "c")]))
                ref = orc.hash_aggregate([b], [0], [(1, "SUM"), (1, "COUNT")])
                assert r.column(0).to_pylist() == [b.column(0)[i].as_py() for i in ref["first_row"]], ("agg", seed)
                assert r.column(1).to_pylist() == ref["f64"][0].tolist(), ("agg", seed)
                assert r.column(2).to_pylist() == ref["i64"][1].tolist(), ("agg", seed)
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
