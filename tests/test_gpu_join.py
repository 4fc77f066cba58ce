"""Join build + probe on the GPU against the oracle (HashTable::build/probe,
hash_join.rs:87-200): pairs bit-exact in the reference's emission order."""
import numpy as np
import pytest

from gpu_util import exact_sum, i64_batch, say, to_dev

import oracle as orc
from prism_executor import _lib as L
from prism_executor import datagen as dg
from prism_executor import device as D

pytestmark = pytest.mark.gpu

KINDS = [L.TABLE_AUTO, L.TABLE_HASHED, L.TABLE_DENSE]


def check_pairs(ctx, bk, pk, kind, pp=None, bp=None, **kw):
    t = D.JoinTable(ctx, to_dev(bk), kind)
    pi, bi, r = t.probe(to_dev(pk), None if pp is None else to_dev(pp), None if bp is None else to_dev(bp), **kw)
    opi, obi = orc.join_probe(i64_batch(bk), [0], i64_batch(pk), [0])
    gpi = pi.numpy().view(np.uint32)
    gbi = bi.numpy().view(np.uint32)
    assert r.matches == len(opi)
    assert np.array_equal(gpi, opi.astype(np.uint32) + np.uint32(kw.get("probe_row_base", 0)))
    assert np.array_equal(gbi, obi)
    if pp is not None:
        assert r.probe_sum == exact_sum(np.asarray(pp)[opi])
    if bp is not None:
        assert r.build_sum == exact_sum(np.asarray(bp, np.int64)[obi])
    return t, r


@pytest.mark.parametrize("kind", KINDS)
def test_unique_permutation_join(ctx, kind):
    say("unique permutation join", kind)
    d = dg.join_inputs(100_000, 1_000_000)
    t, r = check_pairs(ctx, d["build_key"], d["probe_key"], kind, d["probe_payload"], d["build_payload"])
    assert t.info.unique == 1
    if kind == L.TABLE_AUTO:
        assert t.info.kind == L.TABLE_DENSE


@pytest.mark.parametrize("kind", KINDS)
def test_duplicate_keys_zipf(ctx, kind):
    say("duplicate zipf join", kind)
    cdf = dg.zipf_cdf(5_000, 1.1)
    bk = dg.zipf_i64(20_000, 11, 40, cdf, 41)
    pk = dg.zipf_i64(60_000, 11, 43, cdf, 41)
    pp = dg.bits_i64(60_000, 11, 44)
    bp = dg.uniform_i32(20_000, 11, 45, -(1 << 31), 1 << 32)
    t, r = check_pairs(ctx, bk, pk, kind, pp, bp)
    assert t.info.unique == 0 and t.info.max_dup > 100


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dups", [False, True])
def test_payload_in_table_matches_gather(ctx, kind, dups):
    say("payload in table", kind, dups)
    if dups:
        cdf = dg.zipf_cdf(3_000, 1.1)
        bk = dg.zipf_i64(10_000, 12, 40, cdf, 41)
        pk = dg.zipf_i64(40_000, 12, 43, cdf, 41)
    else:
        d = dg.join_inputs(30_000, 200_000)
        bk, pk = d["build_key"], d["probe_key"]
    bp = dg.uniform_i32(len(bk), 12, 45, -(1 << 31), 1 << 32)
    t = D.JoinTable(ctx, to_dev(bk), kind, payload=to_dev(bp))
    pi, bi, r = t.probe(to_dev(pk))
    opi, obi = orc.join_probe(i64_batch(bk), [0], i64_batch(pk), [0])
    assert np.array_equal(pi.numpy(), opi) and np.array_equal(bi.numpy(), obi)
    assert r.build_sum == exact_sum(bp.astype(np.int64)[obi])
    assert r.probe_sum == 0


def test_sparse_extreme_keys_hashed(ctx):
    say("extreme keys")
    rng = np.random.default_rng(3)
    bk = np.concatenate([rng.integers(-(2**63), 2**63 - 1, 5000, dtype=np.int64),
                         np.array([-(2**63), 2**63 - 1, 0, -1, -(2**63)], np.int64)])
    pk = np.concatenate([bk[::3], rng.integers(-(2**63), 2**63 - 1, 5000, dtype=np.int64),
                         np.array([-(2**63), 0, 7], np.int64)])
    t, _ = check_pairs(ctx, bk, pk, L.TABLE_AUTO)
    assert t.info.kind == L.TABLE_HASHED


def test_empty_sides(ctx):
    say("empty sides")
    check_pairs(ctx, np.zeros(0, np.int64), np.arange(100, dtype=np.int64), L.TABLE_AUTO)
    check_pairs(ctx, np.arange(100, dtype=np.int64), np.zeros(0, np.int64), L.TABLE_AUTO)


def test_capacity_retry_and_row_base(ctx):
    say("capacity retry")
    d = dg.join_inputs(1000, 50_000)
    check_pairs(ctx, d["build_key"], d["probe_key"], L.TABLE_HASHED, capacity=7, probe_row_base=12345)


def test_unaligned_probe_keys(ctx):
    say("unaligned probe keys")
    d = dg.join_inputs(10_000, 100_001)
    bk = to_dev(d["build_key"])
    pk = to_dev(d["probe_key"])[1:]  # 8-byte aligned, not 16: scalar-load path
    t = D.JoinTable(ctx, bk)
    pi, bi, r = t.probe(pk)
    opi, obi = orc.join_probe(i64_batch(d["build_key"]), [0], i64_batch(d["probe_key"][1:]), [0])
    assert np.array_equal(pi.numpy().view(np.uint32), opi)
    assert np.array_equal(bi.numpy().view(np.uint32), obi)


def test_device_datagen_matches_numpy_twin(ctx):
    say("device datagen")
    n = 1 << 20
    x = D.empty(ctx, n, np.int64)
    L.check(ctx.lib.prism_datagen_permutation_i64(ctx.handle, x.ptr, n, dg.SEED, 10))
    ctx.sync()
    assert np.array_equal(x.numpy(), dg.permutation_i64(n, dg.SEED, 10))
    L.check(ctx.lib.prism_datagen_uniform_i64(ctx.handle, x.ptr, n, dg.SEED, 12, 0, 2_000_000))
    ctx.sync()
    assert np.array_equal(x.numpy(), dg.uniform_i64(n, dg.SEED, 12, 0, 2_000_000))
    L.check(ctx.lib.prism_datagen_bits_i64(ctx.handle, x.ptr, n, dg.SEED, 13))
    ctx.sync()
    assert np.array_equal(x.numpy(), dg.bits_i64(n, dg.SEED, 13))
    cdf = dg.zipf_cdf(1 << 16, 1.1)
    dcdf = to_dev(cdf)
    L.check(ctx.lib.prism_datagen_zipf_i64(ctx.handle, x.ptr, n, dg.SEED, 30, dcdf.ptr, len(cdf), 31))
    ctx.sync()
    assert np.array_equal(x.numpy(), dg.zipf_i64(n, dg.SEED, 30, cdf, 31))


def test_full_size_probe_properties(ctx):
    """1M-row build, 64M-row probe: size-independent checks (pair validity,
    order, closed-form count and exact payload sums)."""
    say("full size 1M x 64M")
    nb, npr = 1_000_000, 64_000_000
    bk = D.empty(ctx, nb, np.int64)
    bp = D.empty(ctx, nb, np.int32)
    pk = D.empty(ctx, npr, np.int64)
    pp = D.empty(ctx, npr, np.int64)
    lib = ctx.lib
    L.check(lib.prism_datagen_permutation_i64(ctx.handle, bk.ptr, nb, dg.SEED, 10))
    L.check(lib.prism_datagen_uniform_i32(ctx.handle, bp.ptr, nb, dg.SEED, 11, -(1 << 31), 1 << 32))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, pk.ptr, npr, dg.SEED, 12, 0, 2 * nb))
    L.check(lib.prism_datagen_bits_i64(ctx.handle, pp.ptr, npr, dg.SEED, 13))
    ctx.sync()
    t = D.JoinTable(ctx, bk)
    pi, bi, r = t.probe(pk, pp, bp)
    hk, hpk, hpp, hbp = bk.numpy(), pk.numpy(), pp.numpy(), bp.numpy()
    gpi, gbi = pi.numpy().astype(np.int64), bi.numpy().astype(np.int64)
    mask = hpk < nb
    assert r.matches == int(mask.sum())
    assert np.array_equal(gpi, np.nonzero(mask)[0])
    assert np.array_equal(hk[gbi], hpk[gpi])
    assert r.probe_sum == exact_sum(hpp[mask])
    assert r.build_sum == int(hbp[gbi].astype(np.int64).sum())


@pytest.mark.parametrize("form", ["one_pass", "staged", "stall"])
@pytest.mark.parametrize("kind", [L.TABLE_DENSE, L.TABLE_HASHED])
def test_unique_probe_forms(ctx, monkeypatch, form, kind):
    """Unique tables probed directly: the staged form (lookup pass, scan, move
    pass), the one-pass probe (pairs written at offsets found by looking back
    over the tiles' counts), and the one-pass probe with a look-back that
    counts an unpublished predecessor after a single poll; every form equals
    the oracle, ragged last tile and payload table included."""
    say("unique probe form", form, kind)
    if form in ("one_pass", "stall"):
        monkeypatch.setenv("PRISM_PROBE_ONE_PASS", "1")
    if form == "stall":
        monkeypatch.setenv("PRISM_LOOKBACK_POLLS", "1")
    d = dg.join_inputs(200_000, 3_000_001)
    bk, bp = d["build_key"], d["build_payload"]
    pk, pp = d["probe_key"], d["probe_payload"]
    check_pairs(ctx, bk, pk, kind, pp, bp)
    t = D.JoinTable(ctx, to_dev(bk), kind, payload=to_dev(bp))  # payload carried in the table
    pi, bi, r = t.probe(to_dev(pk), to_dev(pp))
    opi, obi = orc.join_probe(i64_batch(bk), [0], i64_batch(pk), [0])
    assert np.array_equal(pi.numpy().view(np.uint32), opi) and np.array_equal(bi.numpy().view(np.uint32), obi)
    assert r.matches == len(opi) and r.probe_sum == exact_sum(pp[opi])
    assert r.build_sum == exact_sum(bp[obi].astype(np.int64))
    _, _, r2 = t.probe(to_dev(pk), to_dev(pp), want_pairs=False)
    assert r2.matches == len(opi) and r2.probe_sum == r.probe_sum and r2.build_sum == r.build_sum
    t.close()
