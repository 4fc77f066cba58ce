"""Pins the oracle beyond the reference's own asserts: Arrow C++ (pyarrow)
kernels implementing the same Arrow semantics, closed forms of the seeded
workloads, and brute force on tiny inputs."""
import math
import re

import numpy as np
import pyarrow as pa
import pyarrow.compute as pc
import pytest

import oracle as orc
from prism_executor import datagen as dg

FX_K = 0x517cc1b727220a95
M64 = (1 << 64) - 1


def test_group_hash_single_int_key_is_the_key():
    # FxHasher starts at 0: state = 0*K + k (hash_aggregate.rs:30-54)
    keys = pa.array([0, 1, -1, 2**62, -(2**63)], pa.int64())
    b = pa.record_batch([keys], names=["k"])
    for r, k in enumerate(keys.to_pylist()):
        assert orc.group_hash(b, [0], r) == k & M64


def test_group_hash_int32_sign_extends():
    # write_i32 -> write_u64(i as u64): sign extension (hash_aggregate.rs:61-64)
    b32 = pa.record_batch([pa.array([-1, -5, 7], pa.int32())], names=["k"])
    b64 = pa.record_batch([pa.array([-1, -5, 7], pa.int64())], names=["k"])
    for r in range(3):
        assert orc.group_hash(b32, [0], r) == orc.group_hash(b64, [0], r)


def test_group_hash_null_equals_zero_quirk():
    # a null key writes u8 0 -> same word as integer 0: null and 0 share a group
    b = pa.record_batch([pa.array([0, None, 3], pa.int64())], names=["k"])
    assert orc.group_hash(b, [0], 0) == orc.group_hash(b, [0], 1)
    r = orc.hash_aggregate([b], [0], [(0, "COUNT")])
    assert r["n_groups"] == 2 and r["i64"][0].tolist() == [2, 1]


def test_group_hash_no_collisions_on_flag_domain():
    # the (flag A, flag B) domain of the fused filter+group-by workload: 6 distinct ids
    a = [chr(c) for c in dg.FLAG_A]
    b_ = [chr(c) for c in dg.FLAG_B]
    pairs = [(x, y) for x in a for y in b_]
    rb = pa.record_batch([pa.array([p[0] for p in pairs]), pa.array([p[1] for p in pairs])], names=["a", "b"])
    hs = {orc.group_hash(rb, [0, 1], r) for r in range(len(pairs))}
    assert len(hs) == len(pairs)


def test_group_hash_utf8_word_padding():
    # FxHasher::write zero-pads the last 8-byte chunk and mixes no length
    # (hash_aggregate.rs:42-49): "a" and "a\0" get one group id in the
    # reference.  The 0xff written after each str separates ("ab","c") from ("a","bc").
    rb = pa.record_batch([pa.array(["a", "a\x00", "ab", "a"]), pa.array(["c", "c", "c", "bc"])], names=["x", "y"])
    assert orc.group_hash(rb, [0], 0) == orc.group_hash(rb, [0], 1)
    assert orc.group_hash(rb, [0, 1], 2) != orc.group_hash(rb, [0, 1], 3)
    # closed form for one str: ((0*K + word("a"))*K + 0xff)
    assert orc.group_hash(rb, [0], 0) == ((0x61 * FX_K + 0xff) & M64)


def _pa_groupby(keys, values):
    t = pa.table({"k": keys, "v": values})
    g = t.group_by("k", use_threads=False).aggregate([("v", "sum"), ("v", "count"), ("v", "min"), ("v", "max")])
    return {k: (s, c, mn, mx) for k, s, c, mn, mx in zip(*[g.column(n).to_pylist() for n in
                                                          ("k", "v_sum", "v_count", "v_min", "v_max")])}


def test_groupby_config0_against_arrow_cpp():
    d = dg.groupby_small(200_000)
    rb = pa.record_batch([pa.array(d["key"]), pa.array(d["value"])], names=["k", "v"])
    r = orc.hash_aggregate([rb], [0], [(1, "SUM"), (1, "COUNT"), (1, "MIN"), (1, "MAX")])
    ref = _pa_groupby(d["key"], d["value"])
    assert r["n_groups"] == len(ref)
    for i in range(r["n_groups"]):
        k = int(d["key"][r["first_row"][i]])
        s, c, mn, mx = ref[k]
        assert r["f64"][0][i] == float(s) and r["sum128"][0][i] == s
        assert r["i64"][1][i] == c
        assert r["f64"][2][i] == float(mn) and r["f64"][3][i] == float(mx)
    # groups in first-appearance order
    _, first = np.unique(d["key"], return_index=True)
    assert r["first_row"].tolist() == sorted(first.tolist())


def test_groupby_multibatch_chunk_merge_keeps_first_appearance():
    d = dg.groupby_small(50_000)
    batches = [pa.record_batch([pa.array(d["key"][i:i + 7_000]), pa.array(d["value"][i:i + 7_000])],
                               names=["k", "v"]) for i in range(0, 50_000, 7_000)]
    for threads in (1, 3, 8):
        r = orc.hash_aggregate(batches, [0], [(1, "SUM")], n_threads=threads)
        glob = [7_000 * fb + fr for fb, fr in zip(r["first_batch"], r["first_row"])]
        _, first = np.unique(d["key"], return_index=True)
        assert glob == sorted(first.tolist())
        assert sum(r["sum128"][0]) == int(d["value"].sum())


def test_float_sum_is_left_to_right_f64():
    # GroupAccumulator sums in row order in f64 (hash_aggregate.rs:135-137)
    vals = [1e16, 1.0, -1e16, 1.0]
    rb = pa.record_batch([pa.array([0, 0, 0, 0]), pa.array(vals)], names=["k", "v"])
    r = orc.hash_aggregate([rb], [0], [(1, "SUM")])
    s = 0.0
    for v in vals:
        s += v
    assert r["f64"][0][0] == s == 1.0 and math.fsum(vals) == 2.0


def test_min_max_ignore_nan_and_global_nulls():
    rb = pa.record_batch([pa.array([1, 1, 1]), pa.array([float("nan"), 2.0, -3.0])], names=["k", "v"])
    r = orc.hash_aggregate([rb], [0], [(1, "MIN"), (1, "MAX")])
    assert r["f64"][0][0] == -3.0 and r["f64"][1][0] == 2.0
    g = pa.record_batch([pa.array([None, 5, None, 2], pa.int64())], names=["v"])
    r = orc.hash_aggregate([g], [], [(0, "COUNT"), (0, "SUM"), (0, "AVG"), (0, "MIN"), (0, "MAX")])
    # global path excludes nulls (hash_aggregate.rs:364-377)
    assert r["i64"][0][0] == 2 and r["f64"][1][0] == 7.0 and r["f64"][2][0] == 3.5
    assert r["f64"][3][0] == 2.0 and r["f64"][4][0] == 5.0


def test_grouped_count_includes_nulls_quirk():
    # grouped COUNT counts rows, nulls included (hash_aggregate.rs:257-269)
    rb = pa.record_batch([pa.array([1, 1, 2]), pa.array([None, 4, None], pa.int64())], names=["k", "v"])
    r = orc.hash_aggregate([rb], [0], [(1, "COUNT"), (1, "SUM")])
    assert r["i64"][0].tolist() == [2, 1]


def test_join_closed_form_permutation():
    nb, npb = 50_000, 400_000
    d = dg.join_inputs(nb, npb)
    b = pa.record_batch([pa.array(d["build_key"])], names=["k"])
    p = pa.record_batch([pa.array(d["probe_key"])], names=["k"])
    pi, bi = orc.join_probe(b, [0], p, [0])
    inv = np.empty(nb, np.int64)
    inv[d["build_key"]] = np.arange(nb)
    exp_p = np.nonzero(d["probe_key"] < nb)[0]
    assert np.array_equal(pi, exp_p)
    assert np.array_equal(bi, inv[d["probe_key"][exp_p]])


def test_join_pairs_against_arrow_cpp_with_duplicates():
    rng = np.random.default_rng(5)
    bk = rng.integers(0, 300, 2_000)
    pk = rng.integers(0, 400, 5_000)
    b = pa.record_batch([pa.array(bk)], names=["k"])
    p = pa.record_batch([pa.array(pk)], names=["k"])
    pi, bi = orc.join_probe(b, [0], p, [0])
    tb = pa.table({"k": bk, "b": np.arange(len(bk))})
    tp = pa.table({"k": pk, "p": np.arange(len(pk))})
    j = tp.join(tb, "k", join_type="inner", use_threads=False)
    ref = sorted(zip(j.column("p").to_pylist(), j.column("b").to_pylist()))
    got = list(zip(pi.tolist(), bi.tolist()))
    assert got == ref  # the reference order is exactly (probe asc, build asc)


def test_join_general_path_multikey_utf8_and_null_semantics():
    b = pa.record_batch([pa.array(["x", "y", None, "x"]), pa.array([1, 2, 3, 1], pa.int32())], names=["s", "i"])
    p = pa.record_batch([pa.array(["x", None, "y", "x"]), pa.array([1, 3, 2, 2], pa.int32())], names=["s", "i"])
    pi, bi = orc.join_probe(b, [0, 1], p, [0, 1])
    # null meets null when the stored values agree (hash 0u8 + keys_equal on values)
    assert list(zip(pi.tolist(), bi.tolist())) == [(0, 0), (0, 3), (1, 2), (2, 1)]


def test_join_int64_fast_path_ignores_validity():
    # hash_join.rs:157-168 reads values() only: a null probe key stored as 0 meets key 0
    b = pa.record_batch([pa.array([0, 5], pa.int64())], names=["k"])
    p = pa.record_batch([pa.array([None, 5], pa.int64())], names=["k"])
    pi, bi = orc.join_probe(b, [0], p, [0])
    stored = np.frombuffer(p.column(0).buffers()[1], np.int64)[0]
    assert stored == 0 and list(zip(pi.tolist(), bi.tolist())) == [(0, 0), (1, 1)]


def _like_regex(pattern: str) -> str:
    out, i = [], 0
    while i < len(pattern):
        c = pattern[i]
        if c == "\\" and i + 1 < len(pattern):
            out.append(re.escape(pattern[i + 1]))
            i += 2
            continue
        out.append(".*" if c == "%" else "." if c == "_" else re.escape(c))
        i += 1
    return "".join(out)


@pytest.mark.parametrize("pattern", ["%abc%", "abc%", "%abc", "a_c", "%a%b%c%", "", "%", "_", "a\\%b%", "%é_"])
def test_like_against_arrow_cpp_and_regex(pattern):
    rng = np.random.default_rng(7)
    alpha = list("abcxyz%_") + ["é"]
    vals = ["".join(rng.choice(alpha, rng.integers(0, 9))) for _ in range(3000)] + ["abc", "", "a%b", "aé"]
    got = orc.like([v.encode() for v in vals], pattern.encode())
    rx = re.compile(_like_regex(pattern), re.DOTALL)
    brute = [rx.fullmatch(v) is not None for v in vals]
    assert got.tolist() == brute
    if "\\" not in pattern:
        arrow = pc.match_like(pa.array(vals), pattern).to_pylist()
        assert got.tolist() == arrow


def test_ilike_ascii():
    vals = ["ABC", "aBc", "xabcx", "ABD"]
    assert orc.like([v.encode() for v in vals], b"%abc%", True).tolist() == [True, True, True, False]
    assert pc.match_like(pa.array(vals), "%abc%", ignore_case=True).to_pylist() == [True, True, True, False]


def test_eager_join_agg_against_materialized_pairs():
    cdf = dg.zipf_cdf(500, 1.1)
    bk = dg.zipf_i64(2_000, 3, 40, cdf, 41)
    attr = dg.uniform_i32(2_000, 3, 42, 0, 50)
    pk = dg.zipf_i64(20_000, 3, 43, cdf, 41)
    pay = dg.uniform_i64(20_000, 3, 44, -10**9, 2 * 10**9)
    cnt, sums = orc.join_agg_eager(bk, attr, pk, pay, 50)
    b = pa.record_batch([pa.array(bk)], names=["k"])
    p = pa.record_batch([pa.array(pk)], names=["k"])
    pi, bi = orc.join_probe(b, [0], p, [0])
    exp_c = np.bincount(attr[bi], minlength=50)
    assert np.array_equal(cnt, exp_c)
    exp_s = [0] * 50
    for a, v in zip(attr[bi].tolist(), pay[pi].tolist()):
        exp_s[a] += v
    assert sums == exp_s


def _brute_pairs(pk, bk):
    """Nested-loop equi-join in the reference's emission order (probe row
    ascending, build row ascending)."""
    return [(p, b) for p in range(len(pk)) for b in range(len(bk)) if pk[p] == bk[b]]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_join_types_against_brute_force(seed):
    rng = np.random.default_rng(seed)
    nb, npr = int(rng.integers(0, 40)), int(rng.integers(0, 60))
    bk, pk = rng.integers(0, 25, nb), rng.integers(0, 30, npr)
    build = pa.record_batch([pa.array(bk), pa.array(np.arange(nb) * 10)], names=["k", "bv"])
    probe = pa.record_batch([pa.array(pk), pa.array(np.arange(npr) * 7)], names=["k", "pv"])
    pairs = _brute_pairs(pk.tolist(), bk.tolist())
    mp = {p for p, _ in pairs}
    mb = {b for _, b in pairs}
    up = [p for p in range(npr) if p not in mp]
    ub = [b for b in range(nb) if b not in mb]

    def rows(r):
        return list(zip(*[r.column(i).to_pylist() for i in range(r.num_columns)])) if r.num_rows else []

    def expect(seq):
        return [(None if p is None else int(pk[p]), None if p is None else 7 * p,
                 None if b is None else int(bk[b]), None if b is None else 10 * b) for p, b in seq]

    assert rows(orc.hash_join(probe, build, orc.INNER, [0], [0])) == expect(pairs)
    if nb:  # Left pads with build row 0, which an empty build side lacks
        assert rows(orc.hash_join(probe, build, orc.LEFT, [0], [0])) == expect(pairs + [(p, 0) for p in up])
    assert rows(orc.hash_join(probe, build, orc.RIGHT, [0], [0])) == expect(pairs + [(None, b) for b in ub])
    assert rows(orc.hash_join(probe, build, orc.FULL, [0], [0])) == expect(
        pairs + [(p, None) for p in up] + [(None, b) for b in ub])
    semi = orc.hash_join(probe, build, orc.LEFT_SEMI, [0], [0])
    anti = orc.hash_join(probe, build, orc.LEFT_ANTI, [0], [0])
    assert semi.column(1).to_pylist() == [7 * p for p in sorted(mp)]
    assert anti.column(1).to_pylist() == [7 * p for p in up]
    # Arrow C++'s outer joins hold the same rows (in another order)
    tp, tb = pa.table({"k": pk, "p": np.arange(npr)}), pa.table({"k2": bk, "b": np.arange(nb)})
    if npr and nb:
        j = tp.join(tb, "k", "k2", join_type="full outer", use_threads=False)
        ref = sorted(zip(j.column("p").to_pylist(), j.column("b").to_pylist()), key=str)
        full = orc.hash_join(probe, build, orc.FULL, [0], [0])
        got = [(None if v is None else v // 7, None if w is None else w // 10)
               for v, w in zip(full.column(1).to_pylist(), full.column(3).to_pylist())]
        assert sorted(got, key=str) == ref


def test_join_chunked_drops_empty_and_runs_right_as_inner():
    build = pa.record_batch([pa.array([1, 2, 2], pa.int64())], names=["k"])
    probes = [pa.record_batch([pa.array(v, pa.int64())], names=["k"]) for v in ([2, 9], [9], [1])]
    inner = orc.hash_join_probe_chunked(probes, build, orc.INNER, [0], [0])
    assert [r.num_rows for r in inner] == [2, 1]
    right = orc.hash_join_probe_chunked(probes, build, orc.RIGHT, [0], [0])
    assert [r.to_pydict() for r in right] == [r.to_pydict() for r in inner]
    left = orc.hash_join_probe_chunked(probes, build, orc.LEFT, [0], [0])
    assert [r.column(1).to_pylist() for r in left] == [[2, 2, 1], [1], [1]]  # row 0 pads: key 1


def _canon(v):
    """DistinctValue of a Float64 (hash_aggregate.rs:466-474): +0 for -0, one NaN."""
    if v != v:
        return "nan"
    return 0.0 if v == 0.0 else v


def test_count_distinct_float_and_utf8_against_python_sets():
    rng = np.random.default_rng(3)
    n = 3000
    pool = [0.0, -0.0, float("nan"), float(np.frombuffer(np.uint64(0x7ff8dead00000001).tobytes(), np.float64)[0]),
            1.5, -1.5, 2.0**60, 5e-324]
    vals = [pool[i] for i in rng.integers(0, len(pool), n)]
    null = rng.random(n) < 0.1
    words = ["", "a", "b", "ab", "ba", "a\x00", "abcdefghij", "abcdefghik"]
    ss = [words[i] for i in rng.integers(0, len(words), n)]
    keys = rng.integers(0, 7, n)
    rb = pa.record_batch([pa.array(keys), pa.array(vals, mask=null), pa.array(ss, mask=rng.random(n) < 0.1)],
                         names=["k", "f", "s"])
    r = orc.hash_aggregate([rb], [0], [(1, "COUNT_DISTINCT"), (2, "COUNT_DISTINCT")])
    fl, sl = rb.column(1).to_pylist(), rb.column(2).to_pylist()
    for gi, fr in enumerate(r["first_row"]):
        g = keys[fr]
        rows = np.nonzero(keys == g)[0]
        assert r["i64"][0][gi] == len({_canon(fl[i]) for i in rows if fl[i] is not None})
        assert r["i64"][1][gi] == len({sl[i] for i in rows if sl[i] is not None})
