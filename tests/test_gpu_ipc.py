"""Arrow IPC stream decoded straight into HBM (prism_ipc_read_device): the device
columns hold the stream's values, and a device group-by over them equals the
oracle's aggregation of the same batches."""
import ctypes

import numpy as np
import pyarrow as pa
import pytest

from gpu_util import say
from test_ipc import sample, stream

import oracle as orc
from prism_executor import _lib as L
from prism_executor import device as D
from prism_executor.ipc import IpcStream

pytestmark = pytest.mark.gpu


def d2h(ctx, ptr, nbytes):
    out = np.zeros(max(nbytes, 1), np.uint8)
    if nbytes:
        L.check(ctx.lib.prism_copy_to_host(ctx.handle, out.ctypes.data, ptr, nbytes))
    return out[:nbytes]


def bits(buf, off, n):
    return np.unpackbits(buf, bitorder="little")[off:off + n].astype(bool)


def test_device_columns_hold_the_stream(ctx):
    say("ipc device columns")
    bs = [sample(3000, 11), sample(1, 12), sample(2000, 13).slice(7, 1500)]
    s = IpcStream(ctx, stream(bs))
    try:
        assert s.n_batches == 3 and [f[0] for f in s.fields] == bs[0].schema.names
        for i, b in enumerate(bs):
            n, cols = s.batch(i)
            assert n == b.num_rows
            for j, (col, arr) in enumerate(zip(cols, b.columns)):
                off = col.offset
                valid = ~np.asarray(arr.is_null()) if arr.null_count else np.ones(n, bool)
                if arr.null_count:
                    assert col.validity
                    assert np.array_equal(bits(d2h(ctx, col.validity, (off + n + 7) // 8), off, n), valid)
                t = arr.type
                if t == pa.utf8():
                    o = d2h(ctx, col.offsets, 4 * (off + n + 1)).view(np.int32)[off:]
                    data = d2h(ctx, col.data, int(o[-1])).tobytes()
                    got = [data[o[r]:o[r + 1]].decode() if valid[r] else None for r in range(n)]
                    assert got == arr.to_pylist()
                elif t == pa.bool_():
                    got = bits(d2h(ctx, col.data, (off + n + 7) // 8), off, n)
                    exp = np.array([bool(x) if x is not None else False for x in arr.to_pylist()])
                    assert np.array_equal(got[valid], exp[valid])
                else:
                    dt = {pa.int32(): np.int32, pa.int64(): np.int64, pa.float64(): np.float64,
                          pa.date32(): np.int32}[t]
                    w = np.dtype(dt).itemsize
                    got = d2h(ctx, col.data, w * (off + n)).view(dt)[off:]
                    exp = arr.cast(pa.int32()) if t == pa.date32() else arr
                    exp = np.asarray(exp.fill_null(0))
                    assert np.array_equal(got[valid], exp[valid])
    finally:
        s.close()


def test_groupby_over_ipc_columns(ctx):
    """The decoded columns feed a device group-by directly (no Arrow import)."""
    say("ipc -> device group-by")
    rng = np.random.default_rng(14)
    bs = [pa.record_batch([pa.array(rng.integers(0, 50, n)), pa.array(rng.integers(-10**9, 10**9, n))],
                          names=["k", "v"]) for n in (10_000, 7_777, 20_000)]
    s = IpcStream(ctx, stream(bs))
    g = D.GroupBy(ctx, [L.TYPE_INT64, L.TYPE_INT64], [0], [[("LOAD_I", 1)]],
                  [("SUM", 0), ("COUNT", 0), ("MIN", 0), ("MAX", 0)])
    try:
        base = 0
        for i in range(s.n_batches):
            n, cols = s.batch(i)
            g.update(cols, n, base)
            base += n
        out = g.read(g.finish())
    finally:
        g.close()
        s.close()
    ref = orc.hash_aggregate(bs, [0], [(1, "SUM"), (1, "COUNT"), (1, "MIN"), (1, "MAX")])
    assert out["first_row"].tolist() == [sum(b.num_rows for b in bs[:fb]) + fr
                                         for fb, fr in zip(ref["first_batch"], ref["first_row"])]
    for a in range(4):
        assert np.array_equal(out["f64"][a], ref["f64"][a])
