"""Operator service on the GPU: every request's result in the batched, overlapped
mix equals the one-request-per-launch operator's result on the same input
(group-by exact words, probe matches and int128 sums, filter counts and
bitmaps), with and without graph replay, under table eviction and from many
submitting threads.  Every test runs with the group-by tile kernels generated
per plan and with the prebuilt interpreting ones (PRISM_JIT=0)."""
import threading

import numpy as np
import pytest

from gpu_util import say, to_dev

from prism_executor import _lib as L
from prism_executor import datagen as dg
from prism_executor import device as D
from prism_executor.service import Service, filter_mask

pytestmark = pytest.mark.gpu

I64, I32, UTF8 = L.TYPE_INT64, L.TYPE_INT32, L.TYPE_UTF8
GB0 = dict(column_types=[I64, I64], keys=[0], values=[[("LOAD_I", 1)]],
           aggs=[("SUM", 0), ("COUNT", 0), ("MIN", 0), ("MAX", 0)])
GB2 = dict(column_types=[I32, I64, I64, UTF8, UTF8], keys=[3, 4],
           values=[[("LOAD_I", 1), ("CONST_I", 0, 100), ("LOAD_I", 2), ("SUB_I",), ("MUL_I",)]],
           aggs=[("SUM", 0), ("COUNT", 0), ("AVG", 0)],
           filter_prog=[("LOAD_I", 0), ("CONST_I", 0, 2470), ("CMP_I", L.CMP["LE"])])
LIKE_ABC = [("LIKE", 0, 0)]


@pytest.fixture(autouse=True, params=["generated", "prebuilt"])
def engine(request, monkeypatch):
    monkeypatch.setenv("PRISM_JIT", "1" if request.param == "generated" else "0")
    say("engine", request.param)
    return request.param


class Pools:
    """Device-resident inputs; requests use slices of them."""

    def __init__(self, ctx, n):
        self.n = n
        d0 = dg.groupby_small(n, seed=3)
        self.k0, self.v0 = to_dev(d0["key"]), to_dev(d0["value"])
        d2 = dg.filter_groupby(n, seed=4)
        self.date, self.price, self.disc = to_dev(d2["date"]), to_dev(d2["price"]), to_dev(d2["discount"])
        self.fa, self.fao = to_dev(d2["flag_a"][0]), to_dev(d2["flag_a"][1])
        self.fb, self.fbo = to_dev(d2["flag_b"][0]), to_dev(d2["flag_b"][1])
        self.pk = to_dev(dg.uniform_i64(n, 5, 1, 0, 400_000))
        self.pp = to_dev(dg.bits_i64(n, 5, 2))
        sb, so = dg.lower_utf8(n, 6, 1, 8, 64)
        self.sb, self.so = to_dev(sb), to_dev(so)

    def gb0_cols(self, off):
        return [D.column(I64, self.k0, offset=off), D.column(I64, self.v0, offset=off)]

    def gb2_cols(self, off):
        return [D.column(I32, self.date, offset=off), D.column(I64, self.price, offset=off),
                D.column(I64, self.disc, offset=off), D.column(UTF8, self.fa, self.fao, offset=off),
                D.column(UTF8, self.fb, self.fbo, offset=off)]

    def str_cols(self, off):
        return [D.column(UTF8, self.sb, self.so, offset=off)]


def tables(n_list, seed=9):
    out = []
    for i, n in enumerate(n_list):
        keys = dg.permutation_i64(n, seed, 50 + i)
        if i % 2:  # sparse keys: hashed table
            keys = keys * 1_000_003 + 17
        out.append((keys, dg.uniform_i32(n, seed, 60 + i, -(1 << 31), 1 << 32)))
    return out


def make_requests(n_pool, count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        n = int(np.exp(rng.uniform(np.log(4096), np.log(n_pool // 2))))
        off = int(rng.integers(0, n_pool - n))
        kind = rng.choice(["gb0", "gb2", "probe", "filter"], p=[0.25, 0.25, 0.4, 0.1])
        reqs.append((kind, n, off, int(rng.integers(0, 3))))
    return reqs


def plain(ctx, P, tabs_dev, kind, n, off, t):
    if kind in ("gb0", "gb2"):
        spec = GB0 if kind == "gb0" else GB2
        g = D.GroupBy(ctx, expected_groups=8, **spec)
        g.update(P.gb0_cols(off) if kind == "gb0" else P.gb2_cols(off), n, 0)
        k = g.finish()
        out = g.read(k)
        g.close()
        return out
    if kind == "probe":
        _, _, r = tabs_dev[t].probe(P.pk[off:off + n], P.pp[off:off + n], None, want_pairs=False)
        return (int(r.matches), r.probe_sum, r.build_sum)
    mask = D.empty(ctx, (n + 63) // 64, np.uint64)
    c = filter_mask(ctx, LIKE_ABC, [UTF8], P.str_cols(off), n, mask, strings=[b"%abc%"])
    return (c, mask.numpy())


def same_groups(a, b):
    return (a["first_row"].tolist() == b["first_row"].tolist() and a["key_hash"].tolist() == b["key_hash"].tolist()
            and a["row_count"].tolist() == b["row_count"].tolist() and a["sum128"] == b["sum128"]
            and all(np.array_equal(x, y) for x, y in zip(a["i64"], b["i64"]))
            and all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a["f64"], b["f64"])))


@pytest.fixture(scope="module")
def world(ctx):
    P = Pools(ctx, 4_000_000)
    tabs = tables([200_000, 300_000, 500_000])
    tabs_dev = [D.JoinTable(ctx, to_dev(k), payload=to_dev(p)) for k, p in tabs]
    return P, tabs, tabs_dev


def run_service(ctx, P, tabs, reqs, **cfg):
    S = Service(ctx, **cfg)
    for i, (k, p) in enumerate(tabs):
        S.register_table(i, k, p)
    S.register_groupby(0, max_groups=1000, **GB0)
    S.register_groupby(2, max_groups=8, **GB2)
    S.register_filter(4, LIKE_ABC, [UTF8], strings=[b"%abc%"])
    masks, tickets = {}, []
    for i, (kind, n, off, t) in enumerate(reqs):
        if kind == "gb0":
            tickets.append(S.submit_groupby(0, P.gb0_cols(off), n))
        elif kind == "gb2":
            tickets.append(S.submit_groupby(2, P.gb2_cols(off), n))
        elif kind == "probe":
            tickets.append(S.submit_probe(t, P.pk[off:off + n], P.pp[off:off + n]))
        else:
            masks[i] = D.empty(ctx, (n + 63) // 64, np.uint64)
            tickets.append(S.submit_filter(4, P.str_cols(off), n, masks[i]))
    out = []
    for i, (tk, (kind, n, off, t)) in enumerate(zip(tickets, reqs)):
        r = S.wait(tk)
        assert r.status == 0, (i, kind, r.status)
        if kind in ("gb0", "gb2"):
            out.append(S.read_groups(tk, 0 if kind == "gb0" else 2, r.count))
        elif kind == "probe":
            out.append((int(r.count), r.probe_sum, r.build_sum))
        else:
            out.append((int(r.count), masks[i].numpy()))
        S.release(tk)
    stats = S.stats()
    S.close()
    return out, stats


def check_all(ctx, P, tabs_dev, reqs, got):
    for i, (req, g) in enumerate(zip(reqs, got)):
        ref = plain(ctx, P, tabs_dev, *req)
        if req[0] in ("gb0", "gb2"):
            assert same_groups(g, ref), (i, req)
        elif req[0] == "probe":
            assert g == ref, (i, req)
        else:
            assert g[0] == ref[0] and np.array_equal(g[1], ref[1]), (i, req)


def test_mix_equals_alone(ctx, world):
    say("service mix vs alone, graphs + tuning")
    P, tabs, tabs_dev = world
    reqs = make_requests(P.n, 60, 1)
    got, stats = run_service(ctx, P, tabs, reqs)
    check_all(ctx, P, tabs_dev, reqs, got)
    assert stats["completed"] == len(reqs)


def test_mix_without_graphs_or_tuning(ctx, world):
    say("service mix, eager launches")
    P, tabs, tabs_dev = world
    reqs = make_requests(P.n, 40, 2)
    got, stats = run_service(ctx, P, tabs, reqs, use_graphs=False, autotune=False)
    check_all(ctx, P, tabs_dev, reqs, got)


def test_graph_replays_match_eager(ctx, world):
    say("service graph replays")
    P, tabs, tabs_dev = world
    # one batch shape over and over: small probes against table 0 only, each
    # batch's few launches replayed from a graph (use_graphs=2: every
    # recurring shape, not only batches of 10 or more launches)
    reqs = [("probe", 5000 + 7 * i, 1000 * i, 0) for i in range(30)]
    S = Service(ctx, autotune=False, max_inflight=1, max_batch_requests=2, use_graphs=2)
    for i, (k, p) in enumerate(tabs):
        S.register_table(i, k, p)
    got = []
    op, ob = D.empty(ctx, 6000, np.uint32), D.empty(ctx, 6000, np.uint32)
    for i, (kind, n, off, t) in enumerate(reqs):
        pairs = i % 2 == 1  # every other request also writes its pairs
        tk = S.submit_probe(t, P.pk[off:off + n], P.pp[off:off + n], probe_idx=op if pairs else None,
                            build_idx=ob if pairs else None, capacity=6000)
        r = S.wait(tk)
        got.append((int(r.count), r.probe_sum, r.build_sum))
        ctx.sync()  # eager work of this process between replays
        pi, bi, rr = tabs_dev[0].probe(P.pk[off:off + n], P.pp[off:off + n], None, want_pairs=True)
        assert got[-1] == (int(rr.matches), rr.probe_sum, rr.build_sum)
        if pairs:
            m = int(rr.matches)
            assert np.array_equal(op.numpy()[:m], pi.numpy()) and np.array_equal(ob.numpy()[:m], bi.numpy())
        S.release(tk)
    op.free()
    ob.free()
    assert S.stats()["graph_launches"] >= 20
    S.close()


def test_lru_eviction_keeps_results_exact(ctx, world):
    say("service eviction")
    P, tabs, tabs_dev = world
    arenas = 2 * (256 << 20)
    budget = arenas + (2 << 20) + 64 * (1 << 20) + 700_000 * 16
    reqs = [("probe", 20_000 + i, 3000 * i, i % 3) for i in range(24)]
    got, stats = run_service(ctx, P, tabs, reqs, hbm_budget=budget, max_inflight=2, autotune=False)
    check_all(ctx, P, tabs_dev, reqs, got)
    assert stats["tables_evicted"] > 0


def test_submit_from_threads(ctx, world):
    say("service threads")
    P, tabs, tabs_dev = world
    S = Service(ctx)
    for i, (k, p) in enumerate(tabs):
        S.register_table(i, k, p)
    S.register_groupby(0, max_groups=1000, **GB0)
    reqs = make_requests(P.n, 48, 3)
    reqs = [r if r[0] in ("probe", "gb0") else ("gb0",) + r[1:] for r in reqs]
    results = [None] * len(reqs)

    def worker(ix):
        for i in ix:
            kind, n, off, t = reqs[i]
            tk = S.submit_groupby(0, P.gb0_cols(off), n) if kind == "gb0" else \
                S.submit_probe(t, P.pk[off:off + n], P.pp[off:off + n])
            r = S.wait(tk)
            results[i] = S.read_groups(tk, 0, r.count) if kind == "gb0" else (int(r.count), r.probe_sum, r.build_sum)
            S.release(tk)

    th = [threading.Thread(target=worker, args=(list(range(k, len(reqs), 4)),)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    S.close()
    check_all(ctx, P, tabs_dev, reqs, results)


def test_errors(ctx, world):
    say("service errors")
    P, tabs, _ = world
    S = Service(ctx)
    with pytest.raises(L.PrismError) as e:
        S.submit_groupby(7, P.gb0_cols(0), 10)
    assert e.value.kind == "InvalidArgument"
    with pytest.raises(L.PrismError) as e:
        S.register_table(1, np.array([1, 2, 2], np.int64))
    assert e.value.kind == "InvalidArgument"
    with pytest.raises(L.PrismError):
        S.register_groupby(0, max_groups=5000, **GB0)
    S.close()


def test_nullable_columns_match_plain(ctx, world):
    say("service nullable columns")
    P, _, _ = world
    rng = np.random.default_rng(5)
    valid = to_dev(np.packbits(rng.random(P.n) > 0.1, bitorder="little"))
    S = Service(ctx)
    S.register_groupby(0, max_groups=1000, **GB0)
    for n, off in [(50_000, 0), (300_000, 123_457), (1_000_000, 64)]:
        cols = [D.column(I64, P.k0, validity=valid, offset=off), D.column(I64, P.v0, validity=valid, offset=off)]
        tk = S.submit_groupby(0, cols, n)
        r = S.wait(tk)
        assert r.status == 0
        got = S.read_groups(tk, 0, r.count)
        S.release(tk)
        g = D.GroupBy(ctx, expected_groups=8, **GB0)
        g.update(cols, n, 0)
        ref = g.read(g.finish())
        g.close()
        assert same_groups(got, ref), (n, off)
    S.close()
