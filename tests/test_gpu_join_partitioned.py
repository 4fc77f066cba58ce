"""HASHED tables cut into LDS-sized partitions and the radix-partitioned probe
(PRISM_PROBE_PARTITIONED): pairs in the reference's order and exact payload
sums equal the oracle's and the direct probe's, for every layout the tuner
tries (one partition held whole in LDS, and 2..2048 partitions), with
sparse, negative and INT64_MIN keys, ragged and empty probes; layouts that
cannot hold the keys are refused; the tuner returns a layout that builds."""
import numpy as np
import pytest

from gpu_util import exact_sum, i64_batch, say, to_dev

import oracle as orc
from prism_executor import _lib as L
from prism_executor import device as D
from prism_executor.context import context

pytestmark = pytest.mark.gpu
I64_MIN = -(1 << 63)


def sides(n_build, n_probe, seed):
    rng = np.random.default_rng(seed)
    bk = np.unique(rng.integers(I64_MIN, (1 << 63) - 1, n_build, dtype=np.int64))
    bk = rng.permutation(np.concatenate([bk, [I64_MIN, -5, 0, 7]]))
    bk = np.unique(bk)  # unique keys, sorted
    bk = rng.permutation(bk)
    hit = bk[rng.integers(0, len(bk), n_probe)]
    miss = rng.integers(I64_MIN, (1 << 63) - 1, n_probe, dtype=np.int64)
    pk = np.where(rng.random(n_probe) < 0.5, hit, miss)
    bp = rng.integers(-(1 << 31), 1 << 31, len(bk)).astype(np.int32)
    pp = rng.integers(-(1 << 62), 1 << 62, n_probe, dtype=np.int64)
    return bk, bp, pk, pp


def expected(bk, bp, pk, pp):
    pi, bi = orc.join_probe(i64_batch(bk), [0], i64_batch(pk), [0])
    return pi, bi, exact_sum(pp[pi]), int(bp[bi].astype(np.int64).sum())


def run(ctx, bk, bp, pk, pp, opts):
    t = D.JoinTable(ctx, to_dev(bk), payload=to_dev(bp), opts=opts)
    pi, bi, r = t.probe(to_dev(pk), to_dev(pp))
    info = t.info
    t.close()
    return pi.numpy(), bi.numpy(), r, info


LAYOUTS = [(256, 0.5), (1024, 0.5), (1024, 0.7), (4096, 0.5), (8192, 0.7)]


@pytest.mark.parametrize("ps,lf", LAYOUTS)
def test_partitioned_probe_matches_oracle(ps, lf):
    say("partitioned probe, slots per partition", ps, "load", lf)
    ctx = context()
    bk, bp, pk, pp = sides(150_000, 1_000_003, ps)
    epi, ebi, eps, ebs = expected(bk, bp, pk, pp)
    opts = L.BuildOpts(L.TABLE_HASHED, L.PROBE_PARTITIONED, ps, lf)
    pi, bi, r, info = run(ctx, bk, bp, pk, pp, opts)
    assert info.probe == L.PROBE_PARTITIONED and info.partition_slots == ps and info.partitions > 1
    assert np.array_equal(pi, epi) and np.array_equal(bi, ebi)
    assert r.matches == len(epi) and r.probe_sum == eps and r.build_sum == ebs
    # the same layout probed directly
    pi2, bi2, r2, _ = run(ctx, bk, bp, pk, pp, L.BuildOpts(L.TABLE_HASHED, L.PROBE_DIRECT, ps, lf))
    assert np.array_equal(pi2, epi) and np.array_equal(bi2, ebi) and r2.probe_sum == eps


def test_single_partition_in_lds_and_ragged_probes():
    say("partitioned probe, one partition in LDS")
    ctx = context()
    for n_probe in (0, 1, 4095, 4097, 70_000):
        bk, bp, pk, pp = sides(1500, n_probe, n_probe + 3)
        epi, ebi, eps, ebs = expected(bk, bp, pk, pp)
        opts = L.BuildOpts(L.TABLE_HASHED, L.PROBE_PARTITIONED, 4096, 0.5)
        pi, bi, r, info = run(ctx, bk, bp, pk, pp, opts)
        assert info.partitions == 1, info.partitions
        assert np.array_equal(pi, epi) and np.array_equal(bi, ebi), n_probe
        assert r.matches == len(epi) and r.probe_sum == eps and r.build_sum == ebs


def test_partitioned_probe_without_payload_and_with_row_base():
    say("partitioned probe, no payloads, row base")
    ctx = context()
    bk, bp, pk, pp = sides(40_000, 300_000, 9)
    t = D.JoinTable(ctx, to_dev(bk), opts=L.BuildOpts(L.TABLE_HASHED, L.PROBE_PARTITIONED, 512, 0.5))
    pi, bi, r = t.probe(to_dev(pk), probe_row_base=1000)
    t.close()
    epi, ebi = orc.join_probe(i64_batch(bk), [0], i64_batch(pk), [0])
    assert np.array_equal(pi.numpy(), epi + 1000) and np.array_equal(bi.numpy(), ebi)
    assert r.probe_sum == 0 and r.build_sum == 0


def test_layouts_that_cannot_hold_the_keys_are_refused():
    say("partitioned layouts refused")
    ctx = context()
    bk = np.unique(np.random.default_rng(5).integers(1, 1 << 62, 100_000))
    with pytest.raises(L.PrismError) as e:  # partitions too small for the keys they would get
        D.JoinTable(ctx, to_dev(bk), opts=L.BuildOpts(L.TABLE_HASHED, L.PROBE_DIRECT, 16, 0.9))
    assert e.value.kind == "InvalidArgument"
    with pytest.raises(L.PrismError) as e:  # duplicates cannot take the partitioned probe
        D.JoinTable(ctx, to_dev(np.array([1, 2, 2, 3] * 1000, np.int64)),
                    opts=L.BuildOpts(L.TABLE_HASHED, L.PROBE_PARTITIONED, 1024, 0.5))
    assert e.value.kind == "InvalidArgument"
    with pytest.raises(L.PrismError):  # LDS holds at most 8192 slots of a partition
        D.JoinTable(ctx, to_dev(bk), opts=L.BuildOpts(L.TABLE_HASHED, L.PROBE_PARTITIONED, 16384, 0.5))


def test_tuner_picks_a_layout_that_builds_and_probes_right():
    say("join layout tuner")
    ctx = context()
    bk, bp, pk, pp = sides(200_000, 2_000_000, 21)
    best, cands = D.JoinTable.tune(ctx, to_dev(bk), to_dev(bp), to_dev(pk[:500_000]), to_dev(pp[:500_000]))
    assert len(cands) >= 5
    ok = [c for c in cands if c.status == 0]
    assert ok and min(c.probe_ms for c in ok) > 0
    fastest = min(ok, key=lambda c: c.probe_ms)
    assert (best.kind, best.probe, best.partition_slots) == (fastest.opts.kind, fastest.opts.probe,
                                                            fastest.opts.partition_slots)
    epi, ebi, eps, ebs = expected(bk, bp, pk, pp)
    pi, bi, r, _ = run(ctx, bk, bp, pk, pp, best)
    assert np.array_equal(pi, epi) and np.array_equal(bi, ebi) and r.build_sum == ebs
