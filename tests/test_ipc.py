"""Arrow IPC streams at the JNI boundary, on the CPU (the host forms need no
device): ingest (prism-jni lib.rs:96-121 decodes the JNI byte array with
arrow's StreamReader), the library's decoder against pyarrow's writer; the
reply (lib.rs:111-121 writes the output with arrow's StreamWriter), the
library's writer against pyarrow's reader and the library's own decoder."""
import ctypes
import datetime
import struct

import numpy as np
import pyarrow as pa
import pytest

from prism_executor import _lib as L
from prism_executor.ipc import read_ipc_host, write_ipc


def stream(batches, schema=None, **opts):
    sink = pa.BufferOutputStream()
    with pa.ipc.new_stream(sink, schema or batches[0].schema, options=pa.ipc.IpcWriteOptions(**opts)) as w:
        for b in batches:
            w.write_batch(b)
    return sink.getvalue().to_pybytes()


def sample(n, seed):
    rng = np.random.default_rng(seed)
    nulls = rng.random(n) < 0.1
    strs = ["".join(chr(97 + c) for c in rng.integers(0, 26, rng.integers(0, 12))) for _ in range(n)]
    return pa.record_batch([
        pa.array(rng.integers(-2**31, 2**31, n).astype(np.int32)),
        pa.array(rng.integers(-2**62, 2**62, n), mask=nulls),
        pa.array(rng.standard_normal(n)),
        pa.array(strs, pa.utf8(), mask=rng.random(n) < 0.05),
        pa.array(rng.random(n) < 0.5, pa.bool_()),
        pa.array([datetime.date(1992, 1, 1) + datetime.timedelta(days=int(d)) for d in rng.integers(0, 2557, n)],
                 pa.date32()),
    ], names=["i32", "i64", "f64", "s", "b", "d"])


def test_round_trip_several_batches():
    bs = [sample(1000, 1), sample(0, 2), sample(4097, 3), sample(1000, 4).slice(13, 500)]
    got, keep = read_ipc_host(stream(bs))
    assert len(got) == len(bs)
    for g, b in zip(got, bs):
        assert g.schema.names == b.schema.names and g.schema.types == b.schema.types
        assert g.equals(b)


def test_nullability_and_schema_only():
    sch = pa.schema([pa.field("k", pa.int64(), nullable=False), pa.field("v", pa.float64())])
    got, keep = read_ipc_host(stream([], schema=sch))
    assert got == []
    b = pa.record_batch([pa.array([1, 2, 3]), pa.array([0.5, None, 2.0])], schema=sch)
    got, keep = read_ipc_host(stream([b]))
    assert got[0].schema.field("k").nullable is False and got[0].schema.field("v").nullable is True
    assert got[0].equals(b)


def status_of(data):
    lib = L.load()
    arr = (L.ArrowArray * 8)()
    n = L.c_uint32()
    sch = L.ArrowSchema()

    st = lib.prism_ipc_read_host(ctypes.cast(ctypes.c_char_p(data), ctypes.c_void_p), len(data), arr, 8,
                                 ctypes.byref(n), ctypes.addressof(sch))
    return st, lib.prism_last_error().decode()


@pytest.mark.parametrize("arr", [pa.array([1.5], pa.float32()), pa.array([1], pa.uint8()),
                                 pa.array([[1, 2]], pa.list_(pa.int64())), pa.array(["x"], pa.large_utf8()),
                                 pa.array(["a", "b", "a"]).dictionary_encode()])
def test_unsupported_fields_are_arrow_errors(arr):
    st, msg = status_of(stream([pa.record_batch([arr], names=["c"])]))
    assert st == 1 and "IPC stream" in msg  # PRISM_ERR_ARROW


def test_truncated_and_compressed_streams_fail():
    data = stream([sample(100, 5)])
    for cut in (3, 20, len(data) // 2, len(data) - 9):
        st, msg = status_of(data[:cut])
        assert st != 0, cut
    try:
        comp = stream([sample(100, 6)], compression="zstd")
    except (pa.ArrowNotImplementedError, ValueError, TypeError):
        pytest.skip("this pyarrow build writes no compressed IPC")
    st, msg = status_of(comp)
    assert st != 0 and "compression" in msg


def forge(data: bytes, old: bytes, new: bytes) -> bytes:
    assert data.count(old) >= 1
    return data.replace(old, new)


def test_forged_row_count_is_refused_without_overflow():
    # 77 rows of one Int64 column without nulls; the batch length and the
    # field node length both become 2^61, whose byte size wraps to 0 in 64 bits
    b = pa.record_batch([pa.array(np.arange(1000, 1077, dtype=np.int64))], names=["k"])
    data = stream([b])
    forged = forge(data, struct.pack("<q", 77), struct.pack("<q", 1 << 61))
    st, msg = status_of(forged)
    assert st == 1 and "too short" in msg, msg  # PRISM_ERR_ARROW


def test_forged_utf8_offsets_are_refused():
    b = pa.record_batch([pa.array(["ab", "cd", "ef"])], names=["s"])
    data = stream([b])
    forged = forge(data, struct.pack("<4i", 0, 2, 4, 6), struct.pack("<4i", 0, 4, 2, 6))
    st, msg = status_of(forged)
    assert st == 1 and "decrease" in msg, msg
    forged = forge(data, struct.pack("<4i", 0, 2, 4, 6), struct.pack("<4i", -4, 2, 4, 6))
    st, msg = status_of(forged)
    assert st == 1 and "outside" in msg, msg


def reply_sample(n, seed):
    """Every type the operators output: the ingest types plus UInt32 row
    indices and exact Decimal128(38,0) sums."""
    b = sample(n, seed)
    rng = np.random.default_rng(seed + 100)
    dec = [None if rng.random() < 0.1 else int(rng.integers(-2**62, 2**62)) * 2**40 + int(rng.integers(0, 2**40))
           for _ in range(n)]
    return pa.record_batch(list(b.columns) + [
        pa.array(rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32), pa.uint32()),
        pa.array(dec, pa.decimal128(38, 0)),
        pa.array(rng.random(n) < 0.3, pa.bool_(), mask=rng.random(n) < 0.2),
    ], names=b.schema.names + ["u32", "dec", "bn"])


def test_reply_stream_reads_back_in_pyarrow():
    bs = [reply_sample(1000, 5), reply_sample(0, 6), reply_sample(3001, 7), reply_sample(777, 8).slice(13, 500),
          reply_sample(64, 9).slice(3, 0)]
    data = write_ipc(bs)
    assert len(data) % 8 == 0
    got = pa.ipc.open_stream(data).read_all()
    assert got.schema.equals(bs[0].schema)
    assert got.equals(pa.Table.from_batches(bs))
    assert [b.num_rows for b in pa.ipc.open_stream(data)] == [b.num_rows for b in bs]


def test_reply_stream_reads_back_in_the_library():
    bs = [sample(2000, 11), sample(999, 12).slice(7, 900)]
    back, _ = read_ipc_host(write_ipc(bs))
    assert len(back) == 2
    for a, b in zip(back, bs):
        assert a.equals(b)


def test_reply_stream_edge_cases():
    assert write_ipc([]) == b""  # no output batch: no stream, as the reference
    schema = pa.schema([pa.field("k", pa.int64(), nullable=False), pa.field("s", pa.utf8())])
    b = pa.record_batch([pa.array([1, 2, 3]), pa.array([None, None, None], pa.utf8())], schema=schema)
    got = pa.ipc.open_stream(write_ipc([b])).read_all()
    assert got.schema.equals(schema) and got.to_pylist() == b.to_pylist()
    assert got.schema.field("k").nullable is False and got.column("s").null_count == 3
    with pytest.raises(ValueError):
        write_ipc([b, sample(3, 1)])
