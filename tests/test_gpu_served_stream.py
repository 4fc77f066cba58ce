"""The served request stream at its stated shapes, every result checked
against the oracle: 8 resident build tables of 1M..32M unique keys (seeded
permutations of 0..n-1, int32 payload), requests of log-uniform 4,096..16M
rows, 50% group-by (the 1M-row aggregation shape and the fused filter +
group-by shape), 40% probes returning their (probe row, build row) pairs, 10%
LIKE '%abc%' over 8..64-byte lowercase Utf8.  Inputs are slices of
device-resident pools; all requests are submitted at once, so batches mix
kinds and sizes and overlap on the device."""
import ctypes

import numpy as np
import pytest

from gpu_util import say

import oracle as orc
from prism_executor import _lib as L
from prism_executor import datagen as dg
from prism_executor import device as D
from prism_executor.service import Service

pytestmark = pytest.mark.gpu

I64, I32, UTF8 = L.TYPE_INT64, L.TYPE_INT32, L.TYPE_UTF8
GB0 = dict(column_types=[I64, I64], keys=[0], values=[[("LOAD_I", 1)]],
           aggs=[("SUM", 0), ("COUNT", 0), ("MIN", 0), ("MAX", 0)])
GB2 = dict(column_types=[I32, I64, I64, UTF8, UTF8], keys=[3, 4],
           values=[[("LOAD_I", 1), ("CONST_I", 0, 100), ("LOAD_I", 2), ("SUB_I",), ("MUL_I",)]],
           aggs=[("SUM", 0), ("COUNT", 0), ("AVG", 0)],
           filter_prog=[("LOAD_I", 0), ("CONST_I", 0, 2470), ("CMP_I", L.CMP["LE"])])
TABLE_MROWS = [1, 2, 3, 5, 8, 12, 20, 32]
MAX_ROWS = 16_000_000
NP = MAX_ROWS + 4096
PROBE_DOMAIN = 64_000_000
S0 = dg.SEED


class World:
    """Pools generated on the device (the library's seeded generators) and
    copied to the host once for the oracle."""

    def __init__(self, ctx):
        lib = ctx.lib
        self.ctx = ctx
        ck = L.check
        self.k0, self.v0 = D.empty(ctx, NP, np.int64), D.empty(ctx, NP, np.int64)
        ck(lib.prism_datagen_uniform_i64(ctx.handle, self.k0.ptr, NP, S0, 1, 0, 1000))
        ck(lib.prism_datagen_uniform_i64(ctx.handle, self.v0.ptr, NP, S0, 2, 0, 1_000_000))
        self.date, self.price, self.disc = D.empty(ctx, NP, np.int32), D.empty(ctx, NP, np.int64), D.empty(ctx, NP, np.int64)
        self.fa, self.fao = D.empty(ctx, NP, np.uint8), D.empty(ctx, NP + 1, np.int32)
        self.fb, self.fbo = D.empty(ctx, NP, np.uint8), D.empty(ctx, NP + 1, np.int32)
        ck(lib.prism_datagen_uniform_i32(ctx.handle, self.date.ptr, NP, S0, 20, 0, 2557))
        ck(lib.prism_datagen_uniform_i64(ctx.handle, self.price.ptr, NP, S0, 21, 90_000, 10_410_000))
        ck(lib.prism_datagen_uniform_i64(ctx.handle, self.disc.ptr, NP, S0, 22, 0, 11))
        ck(lib.prism_datagen_flag_utf8(ctx.handle, self.fa.ptr, self.fao.ptr, NP, S0, 25, dg.FLAG_A, len(dg.FLAG_A)))
        ck(lib.prism_datagen_flag_utf8(ctx.handle, self.fb.ptr, self.fbo.ptr, NP, S0, 26, dg.FLAG_B, len(dg.FLAG_B)))
        self.pk, self.pp = D.empty(ctx, NP, np.int64), D.empty(ctx, NP, np.int64)
        ck(lib.prism_datagen_uniform_i64(ctx.handle, self.pk.ptr, NP, S0, 40, 0, PROBE_DOMAIN))
        ck(lib.prism_datagen_bits_i64(ctx.handle, self.pp.ptr, NP, S0, 41))
        self.sb, self.so = D.empty(ctx, NP * 64, np.uint8), D.empty(ctx, NP + 1, np.int32)
        tb = ctypes.c_uint64()
        ck(lib.prism_datagen_lower_utf8(ctx.handle, self.sb.ptr, NP * 64, self.so.ptr, NP, S0, 42, 8, 64,
                                        ctypes.byref(tb)))
        self.tables = []
        for i, m in enumerate(TABLE_MROWS):
            n = m * 1_000_000
            kk, pl = D.empty(ctx, n, np.int64), D.empty(ctx, n, np.int32)
            ck(lib.prism_datagen_permutation_i64(ctx.handle, kk.ptr, n, S0, 70 + i))
            ck(lib.prism_datagen_uniform_i32(ctx.handle, pl.ptr, n, S0, 80 + i, -(1 << 31), 1 << 32))
            self.tables.append((kk, pl))
        ctx.sync()
        # host copies for the oracle
        self.h = {name: getattr(self, name).numpy() for name in
                  ("k0", "v0", "date", "price", "disc", "fa", "fb", "pk", "pp", "so")}
        self.h["sb"] = self.sb[:int(tb.value)].numpy()
        self.h_tables = [(kk.numpy(), pl.numpy()) for kk, pl in self.tables]
        self.inv = []
        for kk, _ in self.h_tables:
            inv = np.empty(len(kk), np.int64)
            inv[kk] = np.arange(len(kk))
            self.inv.append(inv)

    def gb0_cols(self, o):
        return [D.column(I64, self.k0, offset=o), D.column(I64, self.v0, offset=o)]

    def gb2_cols(self, o):
        return [D.column(I32, self.date, offset=o), D.column(I64, self.price, offset=o),
                D.column(I64, self.disc, offset=o), D.column(UTF8, self.fa, self.fao, offset=o),
                D.column(UTF8, self.fb, self.fbo, offset=o)]

    def str_cols(self, o):
        return [D.column(UTF8, self.sb, self.so, offset=o)]

    def close(self):
        for a in [self.k0, self.v0, self.date, self.price, self.disc, self.fa, self.fao, self.fb, self.fbo, self.pk,
                  self.pp, self.sb, self.so] + [x for t in self.tables for x in t]:
            a.free()


def requests(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(np.exp(rng.uniform(np.log(4096), np.log(MAX_ROWS))))
        off = int(rng.integers(0, NP - n))
        u = rng.random()
        kind = "gb0" if u < 0.25 else "gb2" if u < 0.5 else "probe" if u < 0.9 else "filter"
        out.append((kind, n, off, int(rng.integers(0, len(TABLE_MROWS)))))
    return out


def check_gb0(W, n, off, got):
    h = W.h
    b = orc.Batch([orc.Col(orc.INT64, h["k0"][off:off + n]), orc.Col(orc.INT64, h["v0"][off:off + n])])
    ref = orc.hash_aggregate([b], [0], [(1, "SUM"), (1, "COUNT"), (1, "MIN"), (1, "MAX")])
    assert got["first_row"].tolist() == ref["first_row"].tolist()
    assert got["key_hash"].tolist() == ref["key_hash"].tolist()
    assert got["sum128"][0] == ref["sum128"][0]
    assert got["i64"][1].tolist() == ref["i64"][1].tolist()
    for a in (0, 2, 3):
        assert np.array_equal(got["f64"][a], ref["f64"][a]), a


def check_gb2(W, n, off, got):
    h = W.h
    keep = h["date"][off:off + n] <= 2470
    rows = np.nonzero(keep)[0]
    rev = h["price"][off:off + n][keep] * (100 - h["disc"][off:off + n][keep])
    m = len(rows)
    ones = np.arange(m + 1, dtype=np.int32)
    b = orc.Batch([orc.Col(orc.UTF8, h["fa"][off:off + n][keep], offsets=ones),
                   orc.Col(orc.UTF8, h["fb"][off:off + n][keep], offsets=ones), orc.Col(orc.INT64, rev)])
    ref = orc.hash_aggregate([b], [0, 1], [(2, "SUM"), (2, "COUNT"), (2, "AVG")])
    assert got["first_row"].tolist() == rows[ref["first_row"]].tolist()
    assert got["key_hash"].tolist() == ref["key_hash"].tolist()
    assert got["sum128"][0] == ref["sum128"][0]
    assert got["i64"][1].tolist() == ref["i64"][1].tolist()
    assert np.array_equal(got["f64"][2], ref["f64"][2])  # partial sums stay below 2^53: exact


def check_probe(W, n, off, t, r, p_idx, b_idx):
    """Closed form: table t's keys are a permutation of 0..m-1, so probe key
    k matches iff k < m, with build row perm^-1(k)."""
    keys, pay = W.h["pk"][off:off + n], W.h["pp"][off:off + n]
    bk, bp = W.h_tables[t]
    m, inv = len(bk), W.inv[t]
    hit = np.nonzero(keys < m)[0]
    brow = inv[keys[hit]]
    assert r.count == len(hit)
    assert p_idx.numpy()[:len(hit)].tolist() == hit.tolist()
    assert np.array_equal(b_idx.numpy()[:len(hit)], brow.astype(np.uint32))
    ps = pay[hit]
    assert r.probe_sum == int((ps >> 32).sum()) * (1 << 32) + int((ps & 0xFFFFFFFF).sum())
    assert r.build_sum == int(bp[brow].astype(np.int64).sum())


def check_filter(W, n, off, r, mask):
    h = W.h
    col = orc.Col(orc.UTF8, h["sb"], offsets=h["so"][off:off + n + 1])
    ref = orc.like_column(col, b"%abc%")
    assert r.count == int(ref.sum())
    words = np.packbits(np.concatenate([ref, np.zeros((-n) % 64, bool)]), bitorder="little").view(np.uint64)
    assert np.array_equal(mask.numpy(), words)


@pytest.fixture(scope="module")
def world(ctx):
    W = World(ctx)
    yield W
    W.close()


def test_served_stream_at_stated_shapes_matches_oracle(ctx, world):
    W = world
    reqs = requests(1000, 17)
    say("served stream: 1000 requests, tables", TABLE_MROWS, "M rows, total rows", sum(r[1] for r in reqs))
    S = Service(ctx, hbm_budget=96 << 30, arena_bytes=1 << 30)
    for i, (kk, pl) in enumerate(W.h_tables):
        S.register_table(i, kk, pl)
    S.register_groupby(0, max_groups=1000, **GB0)
    S.register_groupby(2, max_groups=8, **GB2)
    S.register_filter(4, [("LIKE", 0, 0)], [UTF8], strings=[b"%abc%"])
    outs, tickets = {}, []
    for i, (kind, n, off, t) in enumerate(reqs):
        if kind == "gb0":
            tickets.append(S.submit_groupby(0, W.gb0_cols(off), n))
        elif kind == "gb2":
            tickets.append(S.submit_groupby(2, W.gb2_cols(off), n))
        elif kind == "probe":
            outs[i] = (D.empty(ctx, n, np.uint32), D.empty(ctx, n, np.uint32))
            tickets.append(S.submit_probe(t, W.pk[off:off + n], W.pp[off:off + n], probe_idx=outs[i][0],
                                          build_idx=outs[i][1], capacity=n))
        else:
            outs[i] = D.empty(ctx, (n + 63) // 64, np.uint64)
            tickets.append(S.submit_filter(4, W.str_cols(off), n, outs[i]))
    kinds = {}
    for i, (tk, (kind, n, off, t)) in enumerate(zip(tickets, reqs)):
        r = S.wait(tk)
        assert r.status == 0, (i, kind, r.status)
        if kind == "gb0":
            check_gb0(W, n, off, S.read_groups(tk, 0, r.count))
        elif kind == "gb2":
            check_gb2(W, n, off, S.read_groups(tk, 2, r.count))
        elif kind == "probe":
            check_probe(W, n, off, t, r, *outs[i])
            for a in outs.pop(i):
                a.free()
        else:
            check_filter(W, n, off, r, outs[i])
            outs.pop(i).free()
        kinds[kind] = kinds.get(kind, 0) + 1
        S.release(tk)
    st = S.stats()
    S.close()
    say("checked", kinds, "batches", st["batches"], "mean batch requests", round(st["mean_batch_requests"], 2))
    assert st["completed"] == len(reqs) and st["tables_built"] == len(TABLE_MROWS)
    assert st["mean_batch_requests"] > 1  # requests were batched together


def test_served_probe_pairs_respect_capacity(ctx, world):
    """Fewer slots than matches: the first `capacity` pairs in probe order are
    written, the count holds every match (as prism_join_probe_i64 does)."""
    W = world
    S = Service(ctx, hbm_budget=96 << 30)
    kk, pl = W.h_tables[0]
    S.register_table(0, kk, pl)
    n, off, cap = 300_000, 12_345, 1000
    p_idx, b_idx = D.empty(ctx, cap, np.uint32), D.empty(ctx, cap, np.uint32)
    tk = S.submit_probe(0, W.pk[off:off + n], None, probe_idx=p_idx, build_idx=b_idx, capacity=cap)
    r = S.wait(tk)
    S.release(tk)
    S.close()
    keys = W.h["pk"][off:off + n]
    hit = np.nonzero(keys < len(kk))[0]
    assert r.status == 0 and r.count == len(hit) > cap
    inv = W.inv[0]
    assert p_idx.numpy().tolist() == hit[:cap].tolist()
    assert b_idx.numpy().tolist() == inv[keys[hit[:cap]]].tolist()
    p_idx.free()
    b_idx.free()
