"""ctypes wrapper of the oracle (liboracle.so, a CPU restatement of the
reference operators).  Test infrastructure: imported only by tests/, by
__graft_entry__.smoke() as the checker, and by bench.py's CPU-baseline leg.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_int32, c_int64, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

INT32, INT64, FLOAT64, UTF8, BOOL, DATE32 = 1, 2, 3, 4, 5, 6
SUM, COUNT, AVG, MIN, MAX, COUNT_DISTINCT = 0, 1, 2, 3, 4, 5
FUNCS = {"SUM": SUM, "COUNT": COUNT, "AVG": AVG, "MIN": MIN, "MAX": MAX, "COUNT_DISTINCT": COUNT_DISTINCT}


class OrcColumn(Structure):
    _fields_ = [("type", c_int32), ("reserved", c_int32), ("length", c_int64), ("data", c_void_p),
                ("offsets", c_void_p), ("valid", c_void_p)]


class OrcBatch(Structure):
    _fields_ = [("n_rows", c_int64), ("n_columns", c_int32), ("reserved", c_int32),
                ("columns", POINTER(OrcColumn))]


class OrcAgg(Structure):
    _fields_ = [("column", c_int32), ("func", c_int32)]


class OrcAggOut(Structure):
    _fields_ = [("n_groups", c_int64), ("first_batch", POINTER(c_int64)), ("first_row", POINTER(c_int64)),
                ("key_hash", POINTER(c_uint64)), ("f64", POINTER(c_double)), ("i64", POINTER(c_int64)),
                ("sum_lo", POINTER(c_uint64)), ("sum_hi", POINTER(c_int64)), ("cap", c_int64)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built at {LIB_PATH}")
        lib = ctypes.CDLL(LIB_PATH)
        lib.orc_group_hash.restype = c_uint64
        lib.orc_group_hash.argtypes = [POINTER(OrcBatch), POINTER(c_int32), c_int32, c_int64]
        lib.orc_hash_aggregate.restype = c_int32
        lib.orc_hash_aggregate.argtypes = [POINTER(OrcBatch), c_int32, POINTER(c_int32), c_int32,
                                           POINTER(OrcAgg), c_int32, c_int32, POINTER(OrcAggOut)]
        lib.orc_join_probe.restype = c_int64
        lib.orc_join_probe.argtypes = [POINTER(OrcBatch), POINTER(c_int32), POINTER(OrcBatch), POINTER(c_int32),
                                       c_int32, c_void_p, c_void_p, c_int64]
        lib.orc_like.restype = c_int32
        lib.orc_like.argtypes = [c_void_p, c_int32, c_void_p, c_int32, c_int32]
        lib.orc_like_column.restype = None
        lib.orc_like_column.argtypes = [POINTER(OrcColumn), c_void_p, c_int32, c_int32, c_void_p]
        lib.orc_join_agg_eager.restype = c_int32
        lib.orc_join_agg_eager.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_int64,
                                           c_void_p, c_void_p, c_void_p]
        _lib = lib
    return _lib


class Col:
    """A host column for the oracle: numpy values (+ offsets for utf8) and an
    optional per-row validity array.  Keeps the numpy buffers alive."""

    def __init__(self, type_code, data, offsets=None, valid=None):
        self.type = type_code
        self.data = np.ascontiguousarray(data)
        self.offsets = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.int32)
        self.valid = None if valid is None else np.ascontiguousarray(valid, dtype=np.uint8)
        self.length = len(self.offsets) - 1 if self.offsets is not None else len(self.data)

    def c(self) -> OrcColumn:
        return OrcColumn(self.type, 0, self.length, self.data.ctypes.data,
                         None if self.offsets is None else self.offsets.ctypes.data,
                         None if self.valid is None else self.valid.ctypes.data)


def from_arrow(arr) -> Col:
    """pyarrow Array -> Col, reading raw buffers (null slots keep their stored values)."""
    import pyarrow as pa
    t = arr.type
    valid = None
    if arr.null_count:
        valid = np.asarray(arr.is_valid().to_numpy(zero_copy_only=False), dtype=np.uint8)
    bufs = arr.buffers()
    off = arr.offset
    n = len(arr)
    if pa.types.is_string(t):
        offs = np.frombuffer(bufs[1], dtype=np.int32, count=n + 1, offset=off * 4)
        data = np.frombuffer(bufs[2], dtype=np.uint8) if bufs[2] is not None else np.zeros(0, np.uint8)
        return Col(UTF8, data, offsets=offs, valid=valid)
    if pa.types.is_boolean(t):
        bits = np.unpackbits(np.frombuffer(bufs[1], dtype=np.uint8), bitorder="little")[off:off + n]
        return Col(BOOL, bits.astype(np.uint8), valid=valid)
    dt = {pa.int64(): (np.int64, INT64), pa.int32(): (np.int32, INT32), pa.float64(): (np.float64, FLOAT64),
          pa.date32(): (np.int32, DATE32)}[t]
    data = np.frombuffer(bufs[1], dtype=dt[0], count=n, offset=off * np.dtype(dt[0]).itemsize)
    return Col(dt[1], data, valid=valid)


class Batch:
    def __init__(self, cols, n_rows=None):
        self.cols = list(cols)
        self.n_rows = n_rows if n_rows is not None else (self.cols[0].length if self.cols else 0)
        self._arr = (OrcColumn * max(len(self.cols), 1))(*[c.c() for c in self.cols])

    @classmethod
    def from_arrow(cls, rb):
        return cls([from_arrow(rb.column(i)) for i in range(rb.num_columns)], rb.num_rows)

    def c(self) -> OrcBatch:
        return OrcBatch(self.n_rows, len(self.cols), 0, self._arr)


def hash_aggregate(batches, group_by, aggs, n_threads: int = 1) -> dict:
    """hash_aggregate_batches restated (hash_aggregate.rs:182-447).
    aggs: [(column, func_name)].  Returns per-group first (batch,row), key hash,
    f64 / i64 results and exact int128 sums, in output order."""
    lib = load()
    bs = [b if isinstance(b, Batch) else Batch.from_arrow(b) for b in batches]
    arr = (OrcBatch * max(len(bs), 1))(*[b.c() for b in bs])
    total = sum(b.n_rows for b in bs)
    cap = max(total, 1)
    na = len(aggs)
    gb = (c_int32 * max(len(group_by), 1))(*group_by)
    ag = (OrcAgg * max(na, 1))(*[OrcAgg(c, FUNCS[f] if isinstance(f, str) else f) for c, f in aggs])
    fb = np.zeros(cap, np.int64)
    fr = np.zeros(cap, np.int64)
    kh = np.zeros(cap, np.uint64)
    f64 = np.zeros(cap * max(na, 1), np.float64)
    i64 = np.zeros(cap * max(na, 1), np.int64)
    slo = np.zeros(cap * max(na, 1), np.uint64)
    shi = np.zeros(cap * max(na, 1), np.int64)
    P = ctypes.POINTER
    out = OrcAggOut(0, fb.ctypes.data_as(P(c_int64)), fr.ctypes.data_as(P(c_int64)), kh.ctypes.data_as(P(c_uint64)),
                    f64.ctypes.data_as(P(c_double)), i64.ctypes.data_as(P(c_int64)),
                    slo.ctypes.data_as(P(c_uint64)), shi.ctypes.data_as(P(c_int64)), cap)
    rc = lib.orc_hash_aggregate(arr, len(bs), gb, len(group_by), ag, na, n_threads, ctypes.byref(out))
    if rc != 0:
        raise ValueError("oracle: unsupported aggregate input")
    g = int(out.n_groups)
    res = {"n_groups": g, "first_batch": fb[:g].copy(), "first_row": fr[:g].copy(), "key_hash": kh[:g].copy(),
           "f64": [], "i64": [], "sum128": []}
    for a in range(na):
        res["f64"].append(f64[a * cap:a * cap + g].copy())
        res["i64"].append(i64[a * cap:a * cap + g].copy())
        lo = slo[a * cap:a * cap + g]
        hi = shi[a * cap:a * cap + g]
        res["sum128"].append([int(h) * (1 << 64) + int(l) for l, h in zip(lo.tolist(), hi.tolist())])
    return res


def group_hash(batch, key_cols, row) -> int:
    b = batch if isinstance(batch, Batch) else Batch.from_arrow(batch)
    kc = (c_int32 * max(len(key_cols), 1))(*key_cols)
    return int(load().orc_group_hash(ctypes.byref(b.c()), kc, len(key_cols), row))


def join_probe(build, build_keys, probe, probe_keys):
    """HashTable::build + probe restated: (probe_idx, build_idx) uint32 arrays in emission order."""
    lib = load()
    b = build if isinstance(build, Batch) else Batch.from_arrow(build)
    p = probe if isinstance(probe, Batch) else Batch.from_arrow(probe)
    bk = (c_int32 * len(build_keys))(*build_keys)
    pk = (c_int32 * len(probe_keys))(*probe_keys)
    bc, pc = b.c(), p.c()
    cap = max(p.n_rows, 16)
    while True:
        op = np.zeros(cap, np.uint32)
        ob = np.zeros(cap, np.uint32)
        n = lib.orc_join_probe(ctypes.byref(bc), bk, ctypes.byref(pc), pk, len(build_keys), op.ctypes.data,
                               ob.ctypes.data, cap)
        if n <= cap:
            return op[:n], ob[:n]
        cap = n


def like(strings, pattern: bytes, case_insensitive: bool = False) -> np.ndarray:
    """LIKE over a list of bytes objects (string_ops.rs:16-25)."""
    lib = load()
    data = b"".join(strings)
    offs = np.zeros(len(strings) + 1, np.int32)
    offs[1:] = np.cumsum([len(s) for s in strings])
    col = Col(UTF8, np.frombuffer(data, np.uint8) if data else np.zeros(1, np.uint8), offsets=offs)
    out = np.zeros(max(len(strings), 1), np.uint8)
    pat = np.frombuffer(pattern, np.uint8) if pattern else np.zeros(1, np.uint8)
    lib.orc_like_column(ctypes.byref(col.c()), pat.ctypes.data, len(pattern), int(case_insensitive), out.ctypes.data)
    return out[:len(strings)].astype(bool)


def like_column(col: Col, pattern: bytes, case_insensitive: bool = False) -> np.ndarray:
    lib = load()
    out = np.zeros(max(col.length, 1), np.uint8)
    pat = np.frombuffer(pattern, np.uint8) if pattern else np.zeros(1, np.uint8)
    lib.orc_like_column(ctypes.byref(col.c()), pat.ctypes.data, len(pattern), int(case_insensitive), out.ctypes.data)
    return out[:col.length].astype(bool)


def join_agg_eager(build_key, build_attr, probe_key, probe_payload, n_groups):
    """Join -> GROUP BY build attr: (count[g], exact int128 sum[g] of probe payload)."""
    lib = load()
    bk = np.ascontiguousarray(build_key, np.int64)
    ba = np.ascontiguousarray(build_attr, np.int32)
    pk = np.ascontiguousarray(probe_key, np.int64)
    pp = np.ascontiguousarray(probe_payload, np.int64)
    cnt = np.zeros(n_groups, np.uint64)
    lo = np.zeros(n_groups, np.uint64)
    hi = np.zeros(n_groups, np.int64)
    lib.orc_join_agg_eager(bk.ctypes.data, ba.ctypes.data, len(bk), pk.ctypes.data, pp.ctypes.data, len(pk), n_groups,
                           cnt.ctypes.data, lo.ctypes.data, hi.ctypes.data)
    sums = [int(h) * (1 << 64) + int(l) for l, h in zip(lo.tolist(), hi.tolist())]
    return cnt, sums


def expand_worker_aggregates(original):
    """expand_worker_aggregates restated (prism-bench handler.rs:397-428).
    original: [(column, func_name)].  Returns worker [(column, func_name)] and
    per original aggregate its worker indices (AVG -> SUM, COUNT)."""
    worker, mapping = [], []
    for col, f in original:
        if f == "AVG":
            mapping.append([len(worker), len(worker) + 1])
            worker += [(col, "SUM"), (col, "COUNT")]
        else:
            mapping.append([len(worker)])
            worker.append((col, f))
    return worker, mapping


def reducer_merge(partials, group_by_count, original):
    """The reducer of execute_distributed_aggregate restated (handler.rs:311-343):
    build_final_merge_config (430-457) over the partial batches, the merge
    through hash_aggregate_batches, then reconstruct_original_aggregate_batch
    (459-526).  Returns (first (batch, row) of each group, per original
    aggregate its values: int for COUNT / COUNT_DISTINCT, float otherwise)."""
    worker, mapping = expand_worker_aggregates(original)
    merge = [(group_by_count + i, f if f in ("MIN", "MAX") else "SUM") for i, (_, f) in enumerate(worker)]
    r = hash_aggregate(partials, list(range(group_by_count)), merge)
    g = r["n_groups"]
    out = []
    for (col, f), m in zip(original, mapping):
        src = r["f64"][m[0]]
        if f in ("COUNT", "COUNT_DISTINCT"):
            # numeric_value_as_i64 of the merged Float64 sum: `as i64` (truncation)
            out.append([int(v) for v in src.tolist()])
        elif f == "AVG":
            cnt = r["f64"][m[1]]
            out.append([float(s) / float(c) if c > 0 else 0.0 for s, c in zip(src.tolist(), cnt.tolist())])
        else:
            out.append([float(v) for v in src.tolist()])
    return list(zip(r["first_batch"][:g].tolist(), r["first_row"][:g].tolist())), out


# ---------------------------------------------------------------------------
# Join types and joined-batch assembly (pure Python over numpy / pyarrow;
# small inputs only).
# ---------------------------------------------------------------------------
INNER, LEFT, RIGHT, FULL, LEFT_SEMI, LEFT_ANTI = 0, 1, 2, 3, 4, 5


def left_join_indices(probe_rows, mp, mb):
    """left_join_indices (hash_join.rs:455-484): matched pairs in order, then
    every unmatched probe row paired with build row 0 (not null)."""
    matched = np.zeros(probe_rows, bool)
    matched[np.asarray(mp, np.int64)] = True
    un = np.nonzero(~matched)[0]
    p = list(np.asarray(mp).tolist()) + un.tolist()
    b = list(np.asarray(mb).tolist()) + [0] * len(un)
    return p, b


def semi_join_mask(probe_rows, mp):
    """semi_join_mask (hash_join.rs:486-492)."""
    m = np.zeros(probe_rows, bool)
    m[np.asarray(mp, np.int64)] = True
    return m


def right_join_indices(build_rows, mp, mb):
    """right_join_indices (hash_join.rs:494-524): matched pairs, then every
    unmatched build row with a null probe index."""
    matched = np.zeros(build_rows, bool)
    matched[np.asarray(mb, np.int64)] = True
    un = np.nonzero(~matched)[0]
    return list(np.asarray(mp).tolist()) + [None] * len(un), list(np.asarray(mb).tolist()) + un.tolist()


def full_join_indices(probe_rows, build_rows, mp, mb):
    """full_join_indices (hash_join.rs:527-566): matched pairs, unmatched
    probe rows (null build), then unmatched build rows (null probe)."""
    pm = np.zeros(probe_rows, bool)
    bm = np.zeros(build_rows, bool)
    pm[np.asarray(mp, np.int64)] = True
    bm[np.asarray(mb, np.int64)] = True
    up, ub = np.nonzero(~pm)[0].tolist(), np.nonzero(~bm)[0].tolist()
    p = list(np.asarray(mp).tolist()) + up + [None] * len(ub)
    b = list(np.asarray(mb).tolist()) + [None] * len(up) + ub
    return p, b


def assemble_joined_batch(probe, p_idx, build, b_idx, probe_nullable=False, build_nullable=False):
    """assemble_joined_batch[_nullable_probe/_both_nullable] (hash_join.rs:424-453,
    569-625): take of every probe column, then every build column, a build
    name that clashes with a probe name gets "_right"; the nullable forms mark
    the probe (and for FULL the build) fields nullable."""
    import pyarrow as pa
    pi = pa.array(p_idx, pa.uint32())
    bi = pa.array(b_idx, pa.uint32())
    cols, fields = [], []
    for i, f in enumerate(probe.schema):
        fields.append(pa.field(f.name, f.type, True if probe_nullable else f.nullable))
        cols.append(probe.column(i).take(pi))
    names = set(probe.schema.names)
    for i, f in enumerate(build.schema):
        name = f.name + "_right" if f.name in names else f.name
        fields.append(pa.field(name, f.type, True if build_nullable else f.nullable))
        cols.append(build.column(i).take(bi))
    return pa.RecordBatch.from_arrays(cols, schema=pa.schema(fields))


def _filter_probe(probe, mask):
    import pyarrow as pa
    return pa.RecordBatch.from_arrays([probe.column(i).filter(pa.array(mask)) for i in range(probe.num_columns)],
                                      schema=probe.schema)


def hash_join(probe, build, join_type, build_keys, probe_keys):
    """hash_join (hash_join.rs:312-365) over pyarrow RecordBatches."""
    mp, mb = join_probe(build, build_keys, probe, probe_keys)
    npr, nb = probe.num_rows, build.num_rows
    if join_type == INNER:
        return assemble_joined_batch(probe, mp, build, mb)
    if join_type == LEFT:
        p, b = left_join_indices(npr, mp, mb)
        return assemble_joined_batch(probe, p, build, b)
    if join_type == LEFT_SEMI:
        return _filter_probe(probe, semi_join_mask(npr, mp))
    if join_type == LEFT_ANTI:
        return _filter_probe(probe, ~semi_join_mask(npr, mp))
    if join_type == RIGHT:
        p, b = right_join_indices(nb, mp, mb)
        return assemble_joined_batch(probe, p, build, b, probe_nullable=True)
    p, b = full_join_indices(npr, nb, mp, mb)
    return assemble_joined_batch(probe, p, build, b, probe_nullable=True, build_nullable=True)


def hash_join_probe_chunked(probe_batches, build, join_type, build_keys, probe_keys):
    """hash_join_probe_chunked (hash_join.rs:206-282): per probe batch as
    hash_join does, except Right/Full, which fall back to the inner assembly
    (261-270); empty outputs are dropped (274-280)."""
    out = []
    for p in probe_batches:
        jt = INNER if join_type in (RIGHT, FULL) else join_type
        r = hash_join(p, build, jt, build_keys, probe_keys)
        if r.num_rows > 0:
            out.append(r)
    return out
