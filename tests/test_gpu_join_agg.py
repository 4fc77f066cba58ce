"""Join -> GROUP BY build attribute (the fused join-aggregate path,
prism-substrait executor.rs:333-458) on the GPU, checked against the
materialised join (oracle join + brute-force grouping) at small sizes and the
oracle's eager restatement at larger ones."""
import numpy as np
import pyarrow as pa
import pytest

from gpu_util import say, to_dev

import oracle as orc
from prism_executor import datagen as dg
from prism_executor import device as D

pytestmark = pytest.mark.gpu


def materialized(bk, attr, pk, pay):
    """Groups of the materialised join in first-appearance order."""
    b = pa.record_batch([pa.array(bk)], names=["k"])
    p = pa.record_batch([pa.array(pk)], names=["k"])
    pi, bi = orc.join_probe(b, [0], p, [0])
    order, cnt, sums, first = [], {}, {}, {}
    for i, j in zip(pi.tolist(), bi.tolist()):
        a = int(attr[j])
        if a not in cnt:
            order.append(a)
            cnt[a] = 0
            sums[a] = 0
            first[a] = (i, j)
        cnt[a] += 1
        sums[a] += int(pay[i])
    return order, cnt, sums, first


def run(ctx, bk, attr, batches):
    j = D.JoinAggregate(ctx, to_dev(bk), to_dev(attr.astype(np.int32)))
    keep = []
    for pk, pay, base in batches:
        dk, dp = to_dev(pk), (to_dev(pay) if pay is not None else None)
        keep.append((dk, dp))
        j.probe(dk, dp, base)
    n = j.finish()
    out = j.read(n)
    j.close()
    return out


def check(out, order, cnt, sums, first):
    assert out["attr"].tolist() == order
    assert [int(c) for c in out["count"]] == [cnt[a] for a in order]
    assert out["sum"] == [sums[a] for a in order]
    assert list(zip(out["first_probe_row"].tolist(), out["first_build_row"].tolist())) == [first[a] for a in order]


@pytest.mark.parametrize("n_values,n_attr", [(500, 50), (5000, 3)])
def test_small_against_materialized_join(ctx, n_values, n_attr):
    say(f"join-agg small {n_values} {n_attr}")
    d = dg.skewed_join(1_000, 12_000, n_values, n_attr, seed=7)
    out = run(ctx, d["build_key"], d["build_attr"], [(d["probe_key"], d["probe_payload"], 0)])
    check(out, *materialized(d["build_key"], d["build_attr"], d["probe_key"], d["probe_payload"]))


def test_probe_batches_and_negative_payloads(ctx):
    say("join-agg batches")
    d = dg.skewed_join(1_000, 30_000, 800, 40, seed=8)
    pay = d["probe_payload"] - 500_000 + (d["probe_payload"] << 40)
    cuts = [0, 1, 9_999, 20_000, 30_000]
    batches = [(d["probe_key"][a:b], pay[a:b], a) for a, b in zip(cuts[:-1], cuts[1:])]
    out = run(ctx, d["build_key"], d["build_attr"], batches)
    check(out, *materialized(d["build_key"], d["build_attr"], d["probe_key"], pay))


def test_hashed_keys_and_wide_attributes(ctx):
    say("join-agg hashed keys, wide attrs")
    rng = np.random.default_rng(9)
    vals = rng.integers(-(1 << 62), 1 << 62, 700)
    vals[0] = np.iinfo(np.int64).min
    bk = vals[rng.integers(0, 700, 2_500)]
    attr = rng.integers(-(1 << 31), (1 << 31) - 1, 2_500).astype(np.int32)
    attr[::7] = attr[0]
    pk = vals[rng.integers(0, 800, 25_000) % 700]
    pk[::13] = 12345  # misses
    pay = rng.integers(-10**12, 10**12, 25_000)
    out = run(ctx, bk, attr, [(pk, pay, 0)])
    check(out, *materialized(bk, attr, pk, pay))


def test_no_matches_and_no_payload(ctx):
    say("join-agg no matches")
    bk = np.arange(100, dtype=np.int64)
    attr = (np.arange(100) % 7).astype(np.int32)
    out = run(ctx, bk, attr, [(np.arange(1000, 2000, dtype=np.int64), None, 0)])
    assert len(out["attr"]) == 0
    out = run(ctx, bk, attr, [(np.array([5, 5, 13], np.int64), None, 0)])
    assert out["attr"].tolist() == [5, 6] and [int(c) for c in out["count"]] == [2, 1] and out["sum"] == [0, 0]


def test_medium_against_eager_oracle(ctx):
    say("join-agg medium vs eager oracle")
    n_attr = 10_000
    d = dg.skewed_join(200_000, 4_000_000, 100_000, n_attr, seed=10)
    out = run(ctx, d["build_key"], d["build_attr"], [(d["probe_key"], d["probe_payload"], 0)])
    cnt, sums = orc.join_agg_eager(d["build_key"], d["build_attr"], d["probe_key"], d["probe_payload"], n_attr)
    nz = [a for a in range(n_attr) if cnt[a] > 0]
    assert sorted(out["attr"].tolist()) == nz
    for a, c, s in zip(out["attr"].tolist(), out["count"].tolist(), out["sum"]):
        assert c == cnt[a] and s == sums[a]
    fp = out["first_probe_row"]
    assert np.all(np.diff(fp) >= 0)


def eager_check(out, bk, attr, pk, pay, n_attr):
    cnt, sums = orc.join_agg_eager(bk, attr, pk, pay, n_attr)
    nz = [a for a in range(n_attr) if cnt[a] > 0]
    assert sorted(out["attr"].tolist()) == nz
    for a, c, s in zip(out["attr"].tolist(), out["count"].tolist(), out["sum"]):
        assert c == cnt[a] and s == sums[a]
    assert np.all(np.diff(out["first_probe_row"]) >= 0)


@pytest.mark.parametrize("chunk", ["0", "5000", "70001"])
def test_partitioned_tail_chunks_and_overflow(ctx, monkeypatch, chunk):
    """Misses of the LDS cache go through the two-level slot partition: small
    passes (several per probe call) leave level-1 regions only a few hundred
    records deep, so hot partitions overflow into device atomics; the result
    must equal the unpartitioned pass and the eager restatement."""
    say(f"join-agg partitioned tail chunk={chunk}")
    if chunk != "0":
        monkeypatch.setenv("PRISM_JA_CHUNK", chunk)
    d = dg.skewed_join(3_000, 200_000, 60_000, 500, seed=21)
    pay = d["probe_payload"] - 400_000
    pay[::5] = pay[::5] << 30  # some negative, some wide payloads
    batches = [(d["probe_key"][:123_457], pay[:123_457], 0), (d["probe_key"][123_457:], pay[123_457:], 123_457)]
    out = run(ctx, d["build_key"], d["build_attr"], batches)
    monkeypatch.setenv("PRISM_JA_PART", "0")
    ref = run(ctx, d["build_key"], d["build_attr"], batches)
    for k in ("attr", "count", "first_probe_row", "first_build_row"):
        assert np.array_equal(out[k], ref[k]), k
    assert out["sum"] == ref["sum"]
    eager_check(out, d["build_key"], d["build_attr"], d["probe_key"], pay, 500)


@pytest.mark.parametrize("chunk", ["0", "300000"])
def test_partitioned_tail_wide_key_range(ctx, monkeypatch, chunk):
    """A 1M-key range needs several aggregation ranges per level-1 partition
    (fan > 1); uniform probe keys put almost every row in the tail."""
    say(f"join-agg wide key range chunk={chunk}")
    if chunk != "0":
        monkeypatch.setenv("PRISM_JA_CHUNK", chunk)
    rng = np.random.default_rng(22)
    n_attr = 5_000
    bk = rng.integers(0, 1_000_000, 300_000)
    bk[0], bk[1] = 0, 999_999
    attr = rng.integers(0, n_attr, 300_000).astype(np.int32)
    pk = rng.integers(-1_000, 1_001_000, 2_000_000)
    pay = rng.integers(-(1 << 62), 1 << 62, 2_000_000)
    out = run(ctx, bk, attr, [(pk, pay, 0)])
    eager_check(out, bk, attr, pk, pay, n_attr)


@pytest.mark.parametrize("chunk", ["0", "5000"])
def test_partitioned_tail_one_hot_partition(ctx, monkeypatch, chunk):
    """Every probe key falls in the first 4,096 slots of a 60K-slot range while
    the rows hold more distinct keys than the LDS cache: the whole tail lands
    in one level-1 partition, overflowing both the LDS staging buffer of a
    round and (with 5,000-row passes) the partition's level-1 region."""
    say(f"join-agg one hot partition chunk={chunk}")
    if chunk != "0":
        monkeypatch.setenv("PRISM_JA_CHUNK", chunk)
    rng = np.random.default_rng(23)
    bk = np.arange(60_000, dtype=np.int64)[rng.permutation(60_000)]
    bk = np.concatenate([bk, bk[:7_000]])  # duplicate keys: several build rows per key
    attr = rng.integers(0, 300, len(bk)).astype(np.int32)
    pk = rng.integers(0, 4_000, 150_000)
    pay = rng.integers(-10**15, 10**15, 150_000)
    out = run(ctx, bk, attr, [(pk, pay, 0)])
    check(out, *materialized(bk, attr, pk, pay))


@pytest.mark.parametrize("chunk", ["0", "300000"])
def test_unique_build_keys_zipf_probe(ctx, monkeypatch, chunk):
    """Unique build keys give the create-time ranking nothing to go on; the
    probe's Zipf head must then be caught by the cache ways claimed at run time
    (and stay exact either way)."""
    say(f"join-agg unique build, zipf probe chunk={chunk}")
    if chunk != "0":
        monkeypatch.setenv("PRISM_JA_CHUNK", chunk)
    n_vals, n_attr = 50_000, 700
    bk = dg.permutation_i64(n_vals, 31, 5)
    attr = np.random.default_rng(31).integers(0, n_attr, n_vals).astype(np.int32)
    cdf = dg.zipf_cdf(n_vals, 1.1)
    pk = dg.zipf_i64(2_000_000, 32, 6, cdf, 7)
    pay = np.random.default_rng(33).integers(-10**9, 10**9, 2_000_000)
    out = run(ctx, bk, attr, [(pk, pay, 0)])
    eager_check(out, bk, attr, pk, pay, n_attr)


def first_order(bk, attr, pk):
    """Groups of the join in first-appearance order (probe row ascending,
    build row ascending within it) with each group's first pair, vectorised."""
    uk, fi = np.unique(pk, return_index=True)
    pos = np.minimum(np.searchsorted(uk, bk), len(uk) - 1)
    rows = np.nonzero(uk[pos] == bk)[0]
    key = fi[pos[rows]].astype(np.int64) * (1 << 32) + rows
    a = attr[rows]
    o = np.lexsort((key, a))
    a_s, k_s = a[o], key[o]
    head = np.r_[True, a_s[1:] != a_s[:-1]]
    ga, gk = a_s[head], k_s[head]
    order = np.argsort(gk, kind="stable")
    return ga[order], gk[order] >> 32, gk[order] & 0xFFFFFFFF


@pytest.mark.parametrize("shape", ["buckets", "sorted"])
def test_dense_group_folds(ctx, shape):
    """Dense attribute groups are folded per 2,048-value bucket in LDS when the
    range needs at most 4,096 buckets, else over build rows sorted by
    attribute; both give every group's count, exact sum and first pair, in
    first-appearance order."""
    say(f"join-agg dense group fold {shape}")
    rng = np.random.default_rng(41)
    if shape == "buckets":  # 12,000 attribute values, negative ones included: six buckets
        nb, lo_a, hi_a = 400_000, -3_000, 9_000
    else:  # a 9.5M attribute range: more than 4,096 buckets
        nb, lo_a, hi_a = 2_500_000, 0, 9_500_000
    cdf = dg.zipf_cdf(200_000, 1.1)
    bk = dg.zipf_i64(nb, 41, 1, cdf, 2)
    attr = rng.integers(lo_a, hi_a, nb).astype(np.int32)
    attr[0], attr[1] = lo_a, hi_a - 1
    pk = dg.zipf_i64(3_000_000, 41, 3, cdf, 2)
    pay = rng.integers(0, 1_000_000, 3_000_000)
    out = run(ctx, bk, attr, [(pk, pay, 0)])
    ga, gr, gb = first_order(bk, attr, pk)
    assert np.array_equal(out["attr"], ga)
    assert np.array_equal(out["first_probe_row"], gr) and np.array_equal(out["first_build_row"], gb)
    uk, inv, kc = np.unique(pk, return_inverse=True, return_counts=True)
    ks = np.zeros(len(uk), np.int64)
    np.add.at(ks, inv, pay)
    pos = np.minimum(np.searchsorted(uk, bk), len(uk) - 1)
    hit = uk[pos] == bk
    gc = np.zeros(hi_a - lo_a, np.int64)
    gs = np.zeros(hi_a - lo_a, np.int64)
    np.add.at(gc, attr[hit] - lo_a, kc[pos[hit]])
    np.add.at(gs, attr[hit] - lo_a, ks[pos[hit]])
    assert np.array_equal(np.asarray(out["count"], np.int64), gc[ga - lo_a])
    assert np.array_equal(np.asarray(out["sum"], np.int64), gs[ga - lo_a])
