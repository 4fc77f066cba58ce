"""Resident build tables of the operator service under store keys, as
prism-flight's PartitionStore keeps batches under string keys
(shuffle_writer.rs:42-86): put appends rows to a key's table, a probe by the
key's table id sees every row put so far, clear drops one key, drop_exchange
drops every key under "exchange/<id>/", and a put that would repeat a key is
refused with the key keeping its earlier rows."""
import numpy as np
import pytest

from gpu_util import say, to_dev

from prism_executor import _lib as L
from prism_executor import device as D
from prism_executor.service import Service

pytestmark = pytest.mark.gpu


def probe(S, ctx, tid, keys, pay):
    n = len(keys)
    p_idx, b_idx = D.empty(ctx, n, np.uint32), D.empty(ctx, n, np.uint32)
    r = S.wait(S.submit_probe(tid, to_dev(keys), to_dev(pay), probe_idx=p_idx, build_idx=b_idx, capacity=n))
    assert r.status == 0
    out = (int(r.count), p_idx.numpy()[:r.count].copy(), b_idx.numpy()[:r.count].copy(), r.probe_sum, r.build_sum)
    p_idx.free()
    b_idx.free()
    return out


def test_keyed_tables(ctx):
    say("service keyed tables")
    rng = np.random.default_rng(41)
    S = Service(ctx, hbm_budget=8 << 30, arena_bytes=64 << 20)
    ka, kb = rng.permutation(1000).astype(np.int64), rng.permutation(np.arange(1000, 3000)).astype(np.int64)
    pa_, pb = rng.integers(-2**31, 2**31, 1000).astype(np.int32), rng.integers(-2**31, 2**31, 2000).astype(np.int32)
    t1 = S.put_table("exchange/7/build/0", ka, pa_)
    assert t1 >= 2**31 and S.table_id("exchange/7/build/0") == t1
    assert S.put_table("exchange/7/build/0", kb, pb) == t1  # appended: one table of 3000 rows
    t2 = S.put_table("exchange/7/build/1", np.arange(5, dtype=np.int64))
    t3 = S.put_table("exchange/8/build/0", np.arange(50, dtype=np.int64))
    assert len({t1, t2, t3}) == 3

    keys = rng.integers(0, 4000, 20_000).astype(np.int64)
    pay = rng.integers(-2**40, 2**40, 20_000).astype(np.int64)
    build_keys, build_pay = np.concatenate([ka, kb]), np.concatenate([pa_, pb])
    row_of = {int(k): i for i, k in enumerate(build_keys)}
    hit = np.nonzero(keys < 3000)[0]
    brow = np.array([row_of[int(k)] for k in keys[hit]], np.uint32)
    n, pi, bi, ps, bs = probe(S, ctx, t1, keys, pay)
    assert n == len(hit) and np.array_equal(pi, hit.astype(np.uint32)) and np.array_equal(bi, brow)
    assert ps == int(pay[hit].astype(object).sum()) and bs == int(build_pay[brow].astype(np.int64).sum())

    # a batch that repeats a key is refused; the key keeps its 3000 rows
    with pytest.raises(L.PrismError):
        S.put_table("exchange/7/build/0", np.array([5, 4000], np.int64), np.array([1, 2], np.int32))
    assert probe(S, ctx, t1, keys, pay)[0] == len(hit)
    # payload columns must match the key's
    with pytest.raises(L.PrismError):
        S.put_table("exchange/7/build/0", np.array([4001], np.int64))

    S.drop_exchange("7")
    for k in ("exchange/7/build/0", "exchange/7/build/1"):
        with pytest.raises(L.PrismError):
            S.table_id(k)
    with pytest.raises(L.PrismError):
        S.submit_probe(t1, to_dev(keys), None)
    assert S.table_id("exchange/8/build/0") == t3
    assert probe(S, ctx, t3, keys, pay)[0] == int((keys < 50).sum())
    S.clear_table("exchange/8/build/0")
    S.clear_table("no/such/key")  # absent keys: no-op
    with pytest.raises(L.PrismError):
        S.table_id("exchange/8/build/0")
    # numeric ids from 2^31 up belong to keyed tables
    with pytest.raises(L.PrismError):
        S.register_table(2**31, np.arange(3, dtype=np.int64))
    S.close()
