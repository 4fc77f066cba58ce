"""The restatement of the reference's CPU path that the benchmark times as its
CPU baseline (oracle/ref_cpu.c: 64-partition build, batch-parallel probe,
chunked aggregation, the executor's filter -> project -> aggregate chain and
fused join -> aggregate) gives the oracle's results, at several thread
counts and batch sizes."""
import numpy as np
import pytest

import oracle as orc
import ref_cpu as rc
from prism_executor import datagen as dg


@pytest.mark.parametrize("threads,batch", [(1, 1 << 20), (3, 5000), (8, 777)])
def test_join_probe_matches_oracle(threads, batch):
    rng = np.random.default_rng(threads)
    bk = rng.integers(-500, 20_000, 30_000)  # duplicates and negative keys
    pk = rng.integers(-1000, 25_000, 60_000)
    t = rc.JoinTable(bk, threads)
    n, pc, bc = t.probe_chunked(pk, threads, batch)
    t.close()
    exp_n = exp_pc = exp_bc = 0
    for lo in range(0, len(pk), batch):
        pi, bi = orc.join_probe(orc.Batch([orc.Col(orc.INT64, bk)]), [0],
                                orc.Batch([orc.Col(orc.INT64, pk[lo:lo + batch])]), [0])
        exp_n += len(pi)
        exp_pc += int(pi.astype(np.uint64).sum())
        exp_bc += int(bi.astype(np.uint64).sum())
    assert (n, pc, bc) == (exp_n, exp_pc % 2**64, exp_bc % 2**64)


def test_join_config1_closed_form():
    d = dg.join_inputs(100_000, 1_000_000)
    t = rc.JoinTable(d["build_key"], 4)
    n, _, bc = t.probe_chunked(d["probe_key"], 4, 1 << 16)
    t.close()
    inv = np.empty(100_000, np.int64)
    inv[d["build_key"]] = np.arange(100_000)
    hit = d["probe_key"] < 100_000
    assert n == int(hit.sum()) and bc == int(inv[d["probe_key"][hit]].sum())


@pytest.mark.parametrize("threads,batch", [(1, 1 << 20), (4, 30_000), (16, 9_999)])
def test_groupby_matches_oracle_chunking(threads, batch):
    d = dg.groupby_small(200_000)
    got = rc.groupby_i64(d["key"], d["value"], threads, batch)
    bs = [orc.Batch([orc.Col(orc.INT64, d["key"][i:i + batch]), orc.Col(orc.INT64, d["value"][i:i + batch])])
          for i in range(0, 200_000, batch)]
    ref = orc.hash_aggregate(bs, [0], [(1, "SUM"), (1, "COUNT"), (1, "MIN"), (1, "MAX")], n_threads=threads)
    glob = [batch * b + r for b, r in zip(ref["first_batch"], ref["first_row"])]
    assert got["n_groups"] == ref["n_groups"]
    assert got["first_row"] == glob and got["key_hash"] == ref["key_hash"].tolist()
    for a in range(4):
        assert got["agg"][a] == ref["f64"][a].tolist(), a


@pytest.mark.parametrize("threads,batch", [(1, 1 << 20), (3, 20_000), (8, 4_321)])
def test_filter_project_aggregate_matches_oracle(threads, batch):
    n = 150_000
    d = dg.filter_groupby(n)
    got = rc.filter_project_aggregate(d["date"], d["price"], d["discount"], d["flag_a"], d["flag_b"], 2470, threads,
                                      batch)
    # the oracle over the filtered, projected batches (f64 product, as the
    # reference's BinaryOp casts both sides)
    bs = []
    for lo in range(0, n, batch):
        hi = min(n, lo + batch)
        keep = d["date"][lo:hi] <= 2470
        m = int(keep.sum())
        if not m:
            continue
        rev = d["price"][lo:hi][keep].astype(np.float64) * (100.0 - d["discount"][lo:hi][keep].astype(np.float64))
        ones = np.arange(m + 1, dtype=np.int32)
        bs.append(orc.Batch([orc.Col(orc.UTF8, d["flag_a"][0][lo:hi][keep], offsets=ones),
                             orc.Col(orc.UTF8, d["flag_b"][0][lo:hi][keep], offsets=ones),
                             orc.Col(orc.FLOAT64, rev)]))
    ref = orc.hash_aggregate(bs, [0, 1], [(2, "SUM"), (2, "COUNT"), (2, "AVG")], n_threads=threads)
    sizes = np.cumsum([0] + [b.n_rows for b in bs])
    assert got["n_groups"] == ref["n_groups"] == 6
    assert got["first_row"] == [int(sizes[b]) + r for b, r in zip(ref["first_batch"], ref["first_row"])]
    assert got["key_hash"] == ref["key_hash"].tolist()
    for a in range(3):
        assert got["agg"][a] == ref["f64"][a].tolist(), a  # same chunking: bit-identical f64
    # the float64 variant: price * 0.01 * (1 - discount * 0.01), each BinaryOp in f64
    got = rc.filter_project_aggregate(d["date"], d["price"], d["discount"], d["flag_a"], d["flag_b"], 2470, threads,
                                      batch, variant=1)
    bs2 = []
    for lo in range(0, n, batch):
        hi = min(n, lo + batch)
        keep = d["date"][lo:hi] <= 2470
        m = int(keep.sum())
        if not m:
            continue
        p = d["price"][lo:hi][keep].astype(np.float64) * 0.01
        q = 1.0 - d["discount"][lo:hi][keep].astype(np.float64) * 0.01
        ones = np.arange(m + 1, dtype=np.int32)
        bs2.append(orc.Batch([orc.Col(orc.UTF8, d["flag_a"][0][lo:hi][keep], offsets=ones),
                              orc.Col(orc.UTF8, d["flag_b"][0][lo:hi][keep], offsets=ones),
                              orc.Col(orc.FLOAT64, p * q)]))
    ref = orc.hash_aggregate(bs2, [0, 1], [(2, "SUM"), (2, "COUNT"), (2, "AVG")], n_threads=threads)
    for a in range(3):
        assert got["agg"][a] == ref["f64"][a].tolist(), a


def test_join_aggregate_matches_eager_oracle():
    d = dg.skewed_join(20_000, 60_000, 5_000, 700)
    t = rc.JoinTable(d["build_key"], 4)
    pairs, cnt, sm = t.join_aggregate(d["build_attr"], d["probe_key"], d["probe_payload"], 700, 4, 7_000)
    t.close()
    ec, es = orc.join_agg_eager(d["build_key"], d["build_attr"], d["probe_key"], d["probe_payload"], 700)
    assert pairs == int(ec.sum())
    assert cnt.tolist() == ec.astype(np.float64).tolist()
    assert all(abs(s - e) <= 1e-12 * max(abs(e), 1) for s, e in zip(sm.tolist(), es))


def test_host_facts():
    assert 1 <= rc.host_threads() <= 16
    assert rc.cpu_model()
