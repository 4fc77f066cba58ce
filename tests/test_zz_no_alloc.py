"""prism_join_probe_i64_noalloc, the probe that allocates nothing (its scratch
is a fixed device array): through the C ABI its outputs equal
prism_join_probe_i64's on the same inputs, for two calls in a row, the second
with a smaller input (the array keeps the first call's contents), for unique
and duplicate-key tables, with and without pairs.  Unique tables take the
one-pass probe there (no pair staging) and the staged one through the regular
entry, so a 64M-row probe with pairs fits the array as well; a probe too
large for the array is refused before anything is launched."""
import ctypes

import numpy as np
import pytest

from gpu_util import say, to_dev

from prism_executor import _lib as L
from prism_executor import datagen as dg
from prism_executor import device as D
from prism_executor.context import context

pytestmark = pytest.mark.gpu


def probe(fn, t, keys, pay, n, want_pairs, cap):
    ctx = context()
    res = D.empty(ctx, ctypes.sizeof(L.ProbeResult), np.uint8)
    op = D.empty(ctx, max(cap, 1), np.uint32) if want_pairs else None
    ob = D.empty(ctx, max(cap, 1), np.uint32) if want_pairs else None
    L.check(fn(ctx.handle, t.handle, keys.ptr, n, 0, pay.ptr, None, D.ptr(op), D.ptr(ob), cap if want_pairs else 0,
               res.ptr))
    ctx.sync()
    r = L.ProbeResult.from_buffer_copy(res.numpy().tobytes())
    m = min(int(r.matches), cap)
    out = (int(r.matches), r.probe_sum, r.build_sum, r.error,
           op.numpy()[:m].tolist() if want_pairs else None, ob.numpy()[:m].tolist() if want_pairs else None)
    for a in (res, op, ob):
        if a is not None:
            a.free()
    return out


@pytest.mark.parametrize("dup", [False, True])
@pytest.mark.parametrize("want_pairs", [True, False])
def test_noalloc_probe_equals_probe_on_two_calls(dup, want_pairs):
    say("no-alloc probe, duplicates", dup, "pairs", want_pairs)
    ctx = context()
    lib = ctx.lib
    nb = 300_000
    bk = dg.permutation_i64(nb, 5, 1)
    if dup:
        bk = bk // 3  # every key three times
    bp = dg.uniform_i32(nb, 5, 2, -(1 << 31), 1 << 32)
    t = D.JoinTable(ctx, to_dev(bk), payload=to_dev(bp))
    for n in (3_000_000, 70_001):  # the second call is smaller
        pk = to_dev(dg.uniform_i64(n, 7, n % 97, 0, 2 * nb))
        pp = to_dev(dg.bits_i64(n, 7, 3))
        cap = 3 * n if dup else n
        a = probe(lib.prism_join_probe_i64, t, pk, pp, n, want_pairs, cap)
        b = probe(lib.prism_join_probe_i64_noalloc, t, pk, pp, n, want_pairs, cap)
        assert a[3] == 0 and b[3] == 0
        assert a == b, n
        pk.free()
        pp.free()
    t.close()


def test_noalloc_full_size_probe_equals_probe():
    """1M-key permutation table with its payload, 64M probe rows: the pairs
    and sums of the allocation-free probe equal the regular probe's."""
    say("no-alloc probe, 1M x 64M")
    ctx = context()
    lib = ctx.lib
    nb, n = 1_000_000, 64_000_000
    bk, bp = D.empty(ctx, nb, np.int64), D.empty(ctx, nb, np.int32)
    pk, pp = D.empty(ctx, n, np.int64), D.empty(ctx, n, np.int64)
    L.check(lib.prism_datagen_permutation_i64(ctx.handle, bk.ptr, nb, dg.SEED, 10))
    L.check(lib.prism_datagen_uniform_i32(ctx.handle, bp.ptr, nb, dg.SEED, 11, -(1 << 31), 1 << 32))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, pk.ptr, n, dg.SEED, 12, 0, 2 * nb))
    L.check(lib.prism_datagen_bits_i64(ctx.handle, pp.ptr, n, dg.SEED, 13))
    ctx.sync()
    t = D.JoinTable(ctx, bk, payload=bp)
    outs = []
    for fn in (lib.prism_join_probe_i64, lib.prism_join_probe_i64_noalloc):
        res = D.empty(ctx, ctypes.sizeof(L.ProbeResult), np.uint8)
        op, ob = D.empty(ctx, n, np.uint32), D.empty(ctx, n, np.uint32)
        L.check(fn(ctx.handle, t.handle, pk.ptr, n, 0, pp.ptr, None, op.ptr, ob.ptr, n, res.ptr))
        ctx.sync()
        r = L.ProbeResult.from_buffer_copy(res.numpy().tobytes())
        m = int(r.matches)
        outs.append((m, r.probe_sum, r.build_sum, r.error, op.numpy()[:m], ob.numpy()[:m]))
        for a in (res, op, ob):
            a.free()
    a, b = outs
    assert a[3] == 0 and b[3] == 0 and a[:3] == b[:3] and a[0] > 31_000_000
    assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])
    for x in (bk, bp, pk, pp):
        x.free()
    t.close()


def test_noalloc_probe_refuses_oversize_before_launch():
    say("no-alloc probe, oversize")
    ctx = context()
    lib = ctx.lib
    # The partitioned probe stages 18 B per probe row (record key, row in tile,
    # result), so 17M rows need more scratch than the array holds.
    opts = L.BuildOpts(L.TABLE_HASHED, L.PROBE_PARTITIONED, 1024, 0.5)
    t = D.JoinTable(ctx, to_dev(np.arange(100_000, dtype=np.int64) * 7919), opts=opts)
    n = 17_000_000
    keys = D.empty(ctx, n, np.int64)
    res = D.empty(ctx, ctypes.sizeof(L.ProbeResult), np.uint8)
    op, ob = D.empty(ctx, 16, np.uint32), D.empty(ctx, 16, np.uint32)
    st = lib.prism_join_probe_i64_noalloc(ctx.handle, t.handle, keys.ptr, n, 0, None, None, op.ptr, ob.ptr, 16,
                                          res.ptr)
    assert st == 3  # PRISM_ERR_INVALID_ARGUMENT
    assert b"scratch" in lib.prism_last_error()
    for a in (keys, res, op, ob):
        a.free()
    t.close()
