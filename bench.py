#!/usr/bin/env python3
"""Benchmark: inner equi-join on int64 keys on one MI355X.

Workload (one step): build a hash table over 1M build rows (unique keys = a
seeded permutation of 0..1M-1, int32 payload), then probe it with 64M rows
(keys U[0, 2M), ~50% match, int64 payload), writing the matched
(probe_row, build_row) index pairs in the reference's order plus the exact
int128 sums of both payloads over the matches.  Inputs are generated on the
device and stay resident in HBM; the timed region holds build + probe.

Prints one JSON line.  The roofline block covers the dominant kernel, the
probe's lookup/compaction pass ("join_probe"): algorithmic bytes per launch =
16 B per probe row (key + payload read) + 8 B per matched pair (two u32
indices written), divided by that kernel's average duration, measured with
HIP events that the library records on the launch stream around each launch;
roofline_whole_probe divides the same bytes by the whole probe call (lookup,
tile scan and the move of the pairs to their final offsets).
The CPU baseline times the restatement of the reference's CPU path
(oracle/ref_cpu.c: its 64-partition build and batch-parallel probe) on the
1M build rows and the first 16 batches of 1,048,576 probe rows, at every host
core (at most 16) and at one core, 2 warm-ups + the median of 5 runs each.
The other sections (1M-row aggregation, fused filter + group-by with its
float64 variant, skewed join -> aggregation, Parquet scan, served request
stream) carry their own roofline and CPU baseline.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "prism-query_amd"))

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of each dominant kernel on the default workloads, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this script (each counter
# scaled by its reading for a plain 1 GiB device copy: FETCH 2.0, WRITE 1.0),
# recorded in profiles/join_probe_1Mx64M.md and profiles/groupby_filter_200M.md.
# Measured offline, not by this run: they hold for the default sizes only.
TRAFFIC = {"join_probe": 3_319_100_000, "groupby_update": 6_133_500_000}
DEFAULT_SIZES = (1_000_000, 64_000_000, 200_000_000)
METRIC = "rows/s & %HBM peak (join probe, group-by); p99 latency batched vs plain"


def progress(section: str, t_start: float) -> None:
    """One stderr line per finished section (stdout carries only the JSON line)."""
    print(f"[bench] {section} done at {time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)


CPU_LABEL = ("C restatement of the reference CPU path (oracle/ref_cpu.c): 64-partition HashMap<u64, Vec<u32>> "
             "build, probe parallel over 1,048,576-row batches, aggregation in per-thread chunks merged in order")


def _ref_cpu():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_cpu
    return ref_cpu


def cpu_protocol(fn, rows: int, unit: str, sample: str, extra: dict = None) -> dict:
    """Time fn(threads) at every host core and at one core: 2 warm-ups, then
    the median of 5 runs each (the reference's scaling bench, scaling.rs:77-93)."""
    rc = _ref_cpu()
    t_all = rc.host_threads()
    med_all, runs_all = rc.timed(lambda: fn(t_all))
    med_one, runs_one = rc.timed(lambda: fn(1))
    out = {"value": rows / med_all, "unit": unit, "cores": t_all, "kind": "port", "sample": sample,
           "value_1core": rows / med_one, "median_s": med_all, "median_s_1core": med_one,
           "runs_s": runs_all, "runs_s_1core": runs_one, "nproc": os.cpu_count(), "cpu_model": rc.cpu_model(),
           "protocol": "2 warm-ups + median of 5, at every host core (at most 16) and at 1 core",
           "label": CPU_LABEL}
    if extra:
        out.update(extra)
    return out


def cpu_baseline(h_build, h_probe) -> dict:
    """Config 1 on the CPU: HashTable::build over the 1M build keys, then
    probe_matches_chunked over the first 16 batches of the probe stream."""
    rc = _ref_cpu()
    n_probe = len(h_probe)
    res = {}

    def run(threads):
        t = rc.JoinTable(h_build, threads)
        res["pairs"] = t.probe_chunked(h_probe, threads)[0]
        t.close()

    out = cpu_protocol(run, n_probe, "rows/s",
                       f"build {len(h_build)} rows + probe the first {n_probe} rows of the same seeded probe stream "
                       f"({(n_probe + rc.BATCH_ROWS - 1) // rc.BATCH_ROWS} batches of {rc.BATCH_ROWS} rows)")
    out["pairs"] = res["pairs"]
    return out


FILTER_LE_2470 = [("LOAD_I", 0), ("CONST_I", 0, 2470), ("CMP_I", 3)]
REVENUE_EXACT = [("LOAD_I", 1), ("CONST_I", 0, 100), ("LOAD_I", 2), ("SUB_I",), ("MUL_I",)]


REVENUE_F64 = [("LOAD_I", 1), ("I2F",), ("CONST_F", 0, 0, 0.01), ("MUL_F",), ("CONST_F", 0, 0, 1.0), ("LOAD_I", 2),
               ("I2F",), ("CONST_F", 0, 0, 0.01), ("MUL_F",), ("SUB_F",), ("MUL_F",)]


def groupby_host_check(date, price, disc, fa, fb, n: int, chunk: int = 8_000_000) -> dict:
    """Per group (flag A, flag B) over the rows with date <= 2470, from the
    device columns copied back chunk by chunk: the exact integer SUM of
    price * (100 - discount) and COUNT (per-chunk f64 bincounts stay below
    2^53, so they are exact), the reference's own SUM of that product in f64,
    added left to right in row order (hash_aggregate.rs:135-137, one batch),
    and the SUM of price * 0.01 * (1 - discount * 0.01) in long double (the
    f64 variant's check value)."""
    import numpy as np

    exact = [0] * 6
    count = [0] * 6
    l2r = [0.0] * 6
    frac = [np.longdouble(0)] * 6
    amap = np.zeros(256, np.int64)
    bmap = np.zeros(256, np.int64)
    for i, c in enumerate(b"ANR"):
        amap[c] = i
    for i, c in enumerate(b"FO"):
        bmap[c] = i
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        keep = date[lo:hi].numpy() <= 2470
        p = price[lo:hi].numpy()[keep]
        d = disc[lo:hi].numpy()[keep]
        gid = amap[fa[lo:hi].numpy()[keep]] * 2 + bmap[fb[lo:hi].numpy()[keep]]
        rev = p * (100 - d)
        s = np.bincount(gid, weights=rev.astype(np.float64), minlength=6)
        c = np.bincount(gid, minlength=6)
        order = np.argsort(gid, kind="stable")
        seg = np.searchsorted(gid[order], np.arange(7))
        rev_f = rev.astype(np.float64)[order]
        fv = ((p.astype(np.float64) * 0.01) * (1.0 - d.astype(np.float64) * 0.01))[order]
        for g in range(6):
            exact[g] += int(s[g])
            count[g] += int(c[g])
            part = rev_f[seg[g]:seg[g + 1]]
            if len(part):
                l2r[g] = float(np.cumsum(np.concatenate([[l2r[g]], part]))[-1])
            frac[g] += np.sum(fv[seg[g]:seg[g + 1]], dtype=np.longdouble)
    return {"exact": exact, "count": count, "l2r": l2r, "frac": [float(x) for x in frac]}


def bench_groupby(ctx, n: int, steps: int, cpu_sample: int, check: bool = True) -> dict:
    """Fused filter + group-by: date int32 <= 2470, group by (flag A, flag B)
    (one-byte Utf8), SUM(price * (100 - discount)) exact int128, COUNT, AVG;
    and the float64 variant, SUM(price * 0.01 * (1 - discount * 0.01)), COUNT,
    AVG in f64.  Algorithmic bytes: 4 (date) + 8 (price) + 8 (discount) + 2 x
    (4 offset + 1 byte) (flags) = 30 B per row.  Both results are checked at
    full size against host sums of the same columns."""
    import numpy as np

    from prism_executor import _lib as L
    from prism_executor import datagen as dg
    from prism_executor import device as D

    lib = ctx.lib
    date = D.empty(ctx, n, np.int32)
    price = D.empty(ctx, n, np.int64)
    disc = D.empty(ctx, n, np.int64)
    fa, fa_off = D.empty(ctx, n, np.uint8), D.empty(ctx, n + 1, np.int32)
    fb, fb_off = D.empty(ctx, n, np.uint8), D.empty(ctx, n + 1, np.int32)
    s = dg.SEED
    L.check(lib.prism_datagen_uniform_i32(ctx.handle, date.ptr, n, s, 20, 0, 2557))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, price.ptr, n, s, 21, 90_000, 10_500_000 - 90_000))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, disc.ptr, n, s, 22, 0, 11))
    L.check(lib.prism_datagen_flag_utf8(ctx.handle, fa.ptr, fa_off.ptr, n, s, 25, dg.FLAG_A, len(dg.FLAG_A)))
    L.check(lib.prism_datagen_flag_utf8(ctx.handle, fb.ptr, fb_off.ptr, n, s, 26, dg.FLAG_B, len(dg.FLAG_B)))
    ctx.sync()
    types = [L.TYPE_INT32, L.TYPE_INT64, L.TYPE_INT64, L.TYPE_UTF8, L.TYPE_UTF8]
    cols = [D.col_i32(date), D.col_i64(price), D.col_i64(disc), D.col_utf8(fa, fa_off), D.col_utf8(fb, fb_off)]

    def run(value_prog):
        g = D.GroupBy(ctx, types, [3, 4], [value_prog], [("SUM", 0), ("COUNT", 0), ("AVG", 0)],
                      filter_prog=FILTER_LE_2470, expected_groups=8)
        for _ in range(2):
            g.reset()
            g.update(cols, n, 0)
            g.finish()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.reset()
            g.update(cols, n, 0)
            k = g.finish()
        wall = (time.perf_counter() - t0) / steps
        out = g.read(k)
        ctx.kernel_timing(True)
        for _ in range(steps):
            L.check(lib.prism_groupby_update(g.handle, (L.Column * len(cols))(*cols), n, 0))
        tot, cnt = ctx.kernel_time("groupby_update")
        ctx.kernel_timing(False)
        g.close()
        upd_ms = tot / max(cnt, 1)
        ach = n * 30 / (upd_ms * 1e-3) / 1e9
        return out, k, wall, upd_ms, {"bound": "hbm", "kernel": "groupby_update", "achieved": ach,
                                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}

    out, k, wall, upd_ms, roof = run(REVENUE_EXACT)
    roof["traffic"] = TRAFFIC.get("groupby_update") if n == DEFAULT_SIZES[2] else None
    out_f, k_f, wall_f, upd_f, roof_f = run(REVENUE_F64)
    if not check:  # profiling runs: the timed work only
        for a in (date, price, disc, fa, fa_off, fb, fb_off):
            a.free()
        return {"workload": "fused filter + group-by (unchecked profiling run)", "update_ms": upd_ms,
                "roofline": roof, "f64_variant": {"update_ms": upd_f, "roofline": roof_f}}
    hc = groupby_host_check(date, price, disc, fa, fb, n)
    # groups come in first-appearance order; the host sums are indexed by
    # (flag A, flag B) -> map each device group through its first row
    hfa, hfb = fa.numpy(), fb.numpy()
    gidx = ["ANR".index(chr(hfa[r])) * 2 + "FO".index(chr(hfb[r])) for r in out["first_row"].tolist()]
    gidx_f = ["ANR".index(chr(hfa[r])) * 2 + "FO".index(chr(hfb[r])) for r in out_f["first_row"].tolist()]
    int_ok = (sorted(gidx) == list(range(6)) and
              all(out["sum128"][0][i] == hc["exact"][g] for i, g in enumerate(gidx)) and
              all(int(out["i64"][1][i]) == hc["count"][g] for i, g in enumerate(gidx)))
    f64_err = max(abs(out_f["f64"][0][i] - hc["frac"][g]) / abs(hc["frac"][g]) for i, g in enumerate(gidx_f))
    ref_dev = [abs(hc["l2r"][g] - hc["exact"][g]) / hc["exact"][g] for g in range(6)]
    rows = int(out["row_count"].sum())
    res = {"workload": f"fused filter + group-by, {n} rows: date<=2470, group (flagA, flagB), "
                       "SUM(price*(100-discount)) int128, COUNT, AVG",
           "rows": n, "groups": k, "rows_passing_filter": rows, "ms_per_step": wall * 1e3,
           "rows_per_s": n / wall, "update_ms": upd_ms, "roofline": roof,
           "check": {"int128_sums_and_counts_exact": bool(int_ok),
                     "reference_f64_sum_max_rel_deviation": max(ref_dev),
                     "reference_f64_sum_rel_deviation_per_group": ref_dev,
                     "note": "the reference sums price*(100-discount) in f64 left to right; per-group exact sums "
                             "pass 2^53, so its result deviates by the listed relative amounts; the device's SUM "
                             "is the exact int128 value"},
           "f64_variant": {"workload": f"the same filter and groups, SUM(price*0.01*(1-discount*0.01)) in f64, COUNT, "
                                       "AVG", "groups": k_f, "ms_per_step": wall_f * 1e3, "rows_per_s": n / wall_f,
                           "update_ms": upd_f, "roofline": roof_f,
                           "check": {"max_rel_error_vs_long_double_sum": f64_err, "tolerance": 1e-12,
                                     "ok": bool(f64_err <= 1e-12 and k_f == 6)}}}
    if cpu_sample:
        rc = _ref_cpu()
        m = min(cpu_sample, n)
        hd, hp, hdi = date[:m].numpy(), price[:m].numpy(), disc[:m].numpy()
        ha = (fa[:m].numpy(), fa_off[:m + 1].numpy())
        hb = (fb[:m].numpy(), fb_off[:m + 1].numpy())
        got = {}
        for variant, key, expr in ((0, None, "price * (100 - discount)"),
                                   (1, "f64_variant", "price * 0.01 * (1 - discount * 0.01)")):
            def run(threads):
                got["r"] = rc.filter_project_aggregate(hd, hp, hdi, ha, hb, 2470, threads, variant=variant)

            cb = cpu_protocol(
                run, m, "rows/s", f"Filter(date <= 2470) -> Project({expr}) -> Aggregate by (flag A, flag B) over "
                f"the first {m} rows ({(m + rc.BATCH_ROWS - 1) // rc.BATCH_ROWS} batches of {rc.BATCH_ROWS} rows), "
                "f64 arithmetic as the reference")
            cb["groups"] = got["r"]["n_groups"]
            (res if key is None else res[key])["cpu_baseline"] = cb
    for a in (date, price, disc, fa, fa_off, fb, fb_off):
        a.free()
    return res


def bench_join_layouts(ctx, steps: int, bk, bp, pk, pp, label: str, sample: int = 8_000_000) -> dict:
    """The join-layout autotuner on one build side: every candidate layout
    (AUTO table probed directly; HASHED tables of load factor 0.5 / 0.7 probed
    directly; HASHED tables cut into 1024..8192-slot partitions probed
    radix-partitioned through LDS) timed on a probe sample, then the full
    probe (pairs + exact payload sums) timed with the tuned layout and with
    the direct HASHED table beside it.  Inputs resident in HBM."""
    import numpy as np

    from prism_executor import _lib as L
    from prism_executor import device as D

    n_probe = pk.numel()
    m = min(sample, n_probe)
    best, cands = D.JoinTable.tune(ctx, bk, bp, pk[:m], pp[:m])
    res = D.empty(ctx, ctypes.sizeof(L.ProbeResult), np.uint8)
    op, ob = D.empty(ctx, n_probe, np.uint32), D.empty(ctx, n_probe, np.uint32)

    def timed_probe(opts):
        t = D.JoinTable(ctx, bk, payload=bp, opts=opts)
        t.probe_async(pk, res, pp, None, op, ob, capacity=n_probe)
        ctx.timer_begin()
        for _ in range(steps):
            t.probe_async(pk, res, pp, None, op, ob, capacity=n_probe)
        ms = ctx.timer_end() / steps
        r = L.ProbeResult.from_buffer_copy(res.numpy().tobytes())
        info = t.info
        t.close()
        return ms, r, info

    tuned_ms, r_t, info_t = timed_probe(best)
    direct_ms, r_d, _ = timed_probe(L.BuildOpts(L.TABLE_HASHED, L.PROBE_DIRECT, 0, 0.5))
    alg = n_probe * 16 + int(r_t.matches) * 8
    same = (r_t.matches, r_t.probe_sum, r_t.build_sum) == (r_d.matches, r_d.probe_sum, r_d.build_sum)
    for a in (res, op, ob):
        a.free()
    form = {0: "direct", 1: "partitioned"}
    return {"workload": label, "matches": int(r_t.matches), "tuned_equals_direct_hashed": bool(same),
            "tuned": {"kind": {0: "auto", 1: "hashed", 2: "dense"}[best.kind], "probe": form[best.probe],
                      "load_factor": best.load_factor or 0.5, "partition_slots": int(best.partition_slots),
                      "partitions": int(info_t.partitions), "table_kind": "dense" if info_t.kind == 2 else "hashed"},
            "tuned_probe_ms": tuned_ms, "tuned_GBps": alg / (tuned_ms * 1e-3) / 1e9,
            "tuned_hbm_frac": alg / (tuned_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "direct_hashed_probe_ms": direct_ms, "speedup_vs_direct_hashed": direct_ms / tuned_ms,
            "tuner_sample_rows": m,
            "candidates": [{"kind": c.opts.kind, "probe": form[c.opts.probe], "load_factor": c.opts.load_factor,
                            "partition_slots": int(c.opts.partition_slots), "status": int(c.status),
                            "build_ms": c.build_ms, "probe_ms": c.probe_ms} for c in cands]}


def bench_large_hashed_join(ctx, steps: int) -> dict:
    """A build side as large as the Infinity Cache: 8M unique keys spread
    over the whole int64 range (a seeded permutation of 0..8M-1 multiplied by
    an odd constant: sparse, so HASHED; 256 MB of slots at load factor 0.5),
    int32 payload; 64M probe rows, half of them build keys; the layout tuner
    against the direct HASHED table."""
    import numpy as np

    from prism_executor import _lib as L
    from prism_executor import datagen as dg
    from prism_executor import device as D

    lib = ctx.lib
    nb, npr = 8_000_000, 64_000_000
    odd = 0x9E3779B97F4A7C15 - (1 << 64)  # an odd multiplier as int64
    bk, bp = D.empty(ctx, nb, np.int64), D.empty(ctx, nb, np.int32)
    pk, pp = D.empty(ctx, npr, np.int64), D.empty(ctx, npr, np.int64)
    L.check(lib.prism_datagen_permutation_i64(ctx.handle, bk.ptr, nb, dg.SEED, 90))
    L.check(lib.prism_datagen_uniform_i32(ctx.handle, bp.ptr, nb, dg.SEED, 91, -(1 << 31), 1 << 32))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, pk.ptr, npr, dg.SEED, 92, 0, 2 * nb))
    L.check(lib.prism_datagen_bits_i64(ctx.handle, pp.ptr, npr, dg.SEED, 93))
    ctx.sync()
    # sparse keys: multiply on the host side of the device buffers
    hb, hp = bk.numpy(), pk.numpy()
    with np.errstate(over="ignore"):
        bk2 = D.to_device(ctx, hb * np.int64(odd))
        pk2 = D.to_device(ctx, hp * np.int64(odd))
    out = bench_join_layouts(ctx, steps, bk2, bp, pk2, pp,
                             f"inner equi-join, {nb} sparse unique int64 build keys (int32 payload) x {npr} probe rows "
                             "(half match, int64 payload): pairs + exact payload sums")
    for a in (bk, bp, pk, pp, bk2, pk2):
        a.free()
    return out


def bench_groupby_1m(ctx, steps: int, cpu: bool) -> dict:
    """Plain hash aggregation of one 1M-row batch: key int64 U[0, 1000), value
    int64 U[0, 10^6); SUM, COUNT, MIN, MAX by key, checked bit for bit against
    the oracle.  Launch-bound at this size (16 MB read)."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from prism_executor import _lib as L
    from prism_executor import datagen as dg
    from prism_executor import device as D

    n = 1_000_000
    lib = ctx.lib
    k, v = D.empty(ctx, n, np.int64), D.empty(ctx, n, np.int64)
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, k.ptr, n, dg.SEED, 1, 0, 1000))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, v.ptr, n, dg.SEED, 2, 0, 1_000_000))
    ctx.sync()
    spec = dict(column_types=[L.TYPE_INT64, L.TYPE_INT64], keys=[0], values=[[("LOAD_I", 1)]],
                aggs=[("SUM", 0), ("COUNT", 0), ("MIN", 0), ("MAX", 0)])
    cols = [D.col_i64(k), D.col_i64(v)]
    g = D.GroupBy(ctx, expected_groups=1000, **spec)

    def step():
        g.reset()
        g.update(cols, n, 0)
        return g.read(g.finish())

    for _ in range(3):
        out = step()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = step()
    wall = (time.perf_counter() - t0) / steps
    ctx.kernel_timing(True)
    for _ in range(steps):
        L.check(lib.prism_groupby_update(g.handle, (L.Column * 2)(*cols), n, 0))
    tot, cnt = ctx.kernel_time("groupby_update")
    ctx.kernel_timing(False)
    g.close()
    hk, hv = k.numpy(), v.numpy()
    ref = orc.hash_aggregate([orc.Batch([orc.Col(orc.INT64, hk), orc.Col(orc.INT64, hv)])], [0],
                             [(1, "SUM"), (1, "COUNT"), (1, "MIN"), (1, "MAX")])
    ok = (out["first_row"].tolist() == ref["first_row"].tolist() and out["key_hash"].tolist() == ref["key_hash"].tolist()
          and all(np.array_equal(out["f64"][a], ref["f64"][a]) for a in (0, 2, 3))
          and np.array_equal(out["i64"][1], ref["i64"][1]))
    upd_ms = tot / max(cnt, 1)
    res = {"workload": "hash aggregation, one batch of 1000000 rows: key int64 U[0,1000), value int64 U[0,1e6); "
                       "SUM/COUNT/MIN/MAX by key (launch-bound at this size; parity config)",
           "groups": int(len(out["first_row"])), "bit_exact_vs_oracle": bool(ok), "ms_per_step": wall * 1e3,
           "rows_per_s": n / wall, "update_ms": upd_ms, "update_GBps": n * 16 / (upd_ms * 1e-3) / 1e9}
    if cpu:
        rc = _ref_cpu()
        res["cpu_baseline"] = cpu_protocol(lambda threads: rc.groupby_i64(hk, hv, threads), n, "rows/s",
                                           "hash_aggregate over the same 1000000-row batch (one batch: the "
                                           "reference's chunking gives it one core at any thread count)")
    k.free()
    v.free()
    return res


def bench_join_agg(ctx, n_build: int, n_probe: int, n_values: int, n_attr: int, steps: int, cpu_sample: int) -> dict:
    """Skewed join then aggregation by a build attribute: build keys Zipf(1.1)
    over n_values with an int32 attribute U[0, n_attr), probe keys Zipf(1.1)
    over the same values with an int64 payload U[0, 10^6); GROUP BY attribute,
    COUNT(*), SUM(payload) exact.  Algorithmic bytes: 12 B per build row + 16 B
    per probe row + 28 B per output group."""
    import numpy as np

    from prism_executor import _lib as L
    from prism_executor import datagen as dg
    from prism_executor import device as D

    lib = ctx.lib
    s = dg.SEED
    cdf = D.to_device(ctx, dg.zipf_cdf(n_values, 1.1))
    bk, ba = D.empty(ctx, n_build, np.int64), D.empty(ctx, n_build, np.int32)
    pk, pp = D.empty(ctx, n_probe, np.int64), D.empty(ctx, n_probe, np.int64)
    L.check(lib.prism_datagen_zipf_i64(ctx.handle, bk.ptr, n_build, s, 30, cdf.ptr, n_values, 31))
    L.check(lib.prism_datagen_uniform_i32(ctx.handle, ba.ptr, n_build, s, 32, 0, n_attr))
    L.check(lib.prism_datagen_zipf_i64(ctx.handle, pk.ptr, n_probe, s, 33, cdf.ptr, n_values, 31))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, pp.ptr, n_probe, s, 34, 0, 1_000_000))
    ctx.sync()

    def step():
        j = D.JoinAggregate(ctx, bk, ba)
        j.probe(pk, pp, 0)
        n = j.finish()
        return j, n

    j, n = step()
    out = j.read(n)
    j.close()
    # full-size checksum: total pairs and payload over pairs = per-key probe
    # totals summed over build rows
    hk, hb = pk.numpy(), bk.numpy()
    cnt_p = np.bincount(hk, minlength=n_values)
    sum_p = np.bincount(hk, weights=pp.numpy().astype(np.float64), minlength=n_values)  # < 2^53: exact
    exp_pairs = int(cnt_p[hb].astype(np.uint64).sum())
    per_row = sum_p[hb].astype(np.int64)  # each < 2^53; total may pass 2^63, so sum in halves
    exp_sum = int((per_row >> 32).sum()) * (1 << 32) + int((per_row & 0xFFFFFFFF).sum())
    got_pairs = int(out["count"].astype(np.uint64).sum())
    got_sum = sum(out["sum"])
    ok = exp_pairs == got_pairs and exp_sum == got_sum and np.all(np.diff(out["first_probe_row"]) >= 0)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        j, n = step()
        j.close()
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    ctx.kernel_timing(True)
    for _ in range(steps):
        j, n = step()
        j.close()
    kern = {}
    for name in ("join_agg_probe", "join_agg_build"):
        tot, cnt = ctx.kernel_time(name)
        kern[name] = {"avg_us": tot / max(cnt, 1) * 1e3, "launches": cnt}
    ctx.kernel_timing(False)
    alg = n_build * 12 + n_probe * 16 + n * 28
    res = {"workload": f"skewed join then aggregation: build {n_build} rows (Zipf(1.1) keys over {n_values}, int32 "
                       f"attribute U[0,{n_attr})), probe {n_probe} rows (Zipf(1.1) keys, int64 payload); GROUP BY "
                       "attribute COUNT(*), SUM(payload) exact, eager (no pair materialisation)",
           "groups": n, "join_pairs": got_pairs, "checksum_ok": bool(ok), "ms_per_step": wall * 1e3,
           "probe_rows_per_s": n_probe / wall, "kernels": kern,
           "roofline": {"bound": "hbm", "kernel": "whole query", "achieved": alg / wall / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": alg / wall / 1e9 / HBM_PEAK_GBS, "alg_bytes": alg}}
    if cpu_sample:
        # The reference's fused path materialises the join output per probe
        # batch (executor.rs:333-458): ~3e5 pairs per probe row here, so the
        # sample is the prefix of the probe stream holding at most
        # `cpu_sample` pairs.  One batch: the reference parallelises only
        # across batches, so every thread count runs it on one core.
        rc = _ref_cpu()
        bcount = np.bincount(hb, minlength=n_values)
        per_row = bcount[hk[:1_000_000]]
        k = int(np.searchsorted(np.cumsum(per_row), cpu_sample, side="right"))
        k = max(k, 1)
        hp = pp[:k].numpy()
        hba = ba.numpy()
        got = {}

        def run(threads):
            t = rc.JoinTable(hb, threads)
            got["pairs"] = t.join_aggregate(hba, hk[:k], hp, n_attr, threads)[0]
            t.close()

        res["cpu_baseline"] = cpu_protocol(
            run, k, "probe rows/s", f"build all {n_build} rows (64-partition table) + the first {k} probe rows "
            f"({int(per_row[:k].sum())} join pairs, materialised per probe batch then aggregated, as the "
            "reference's fused join -> aggregate runs)")
        res["cpu_baseline"]["pairs"] = got["pairs"]
        res["cpu_baseline"]["pairs_per_s"] = got["pairs"] / res["cpu_baseline"]["median_s"]
    for a in (cdf, bk, ba, pk, pp):
        a.free()
    return res


def bench_parquet(ctx, n: int, steps: int) -> dict:
    """Parquet scan of a table shaped like the filter + group-by input (date i32,
    price i64, discount i64, quantity i32, flag A / flag B utf8), written as the
    reference's data generator writes its files (UNCOMPRESSED, chunk statistics,
    dictionary encoding on, 1,048,576-row row groups).  Timed: footer parse and
    decode of every column chunk from host bytes into HBM (host page walk, H2D,
    device index resolution and gather).  pyarrow's multi-threaded reader of
    the same bytes into host memory is reported beside it."""
    import io

    import numpy as np
    import pyarrow as pa
    import pyarrow.parquet as pq

    from prism_executor.parquet_scan import ParquetFile, ParquetTable

    rng = np.random.default_rng(11)
    flag = lambda vals, k: pa.DictionaryArray.from_arrays(pa.array(rng.integers(0, k, n).astype(np.int32)),
                                                          pa.array(vals)).cast(pa.utf8())
    t = pa.table({"date": pa.array(rng.integers(0, 2557, n).astype(np.int32)).cast(pa.date32()),
                  "price": pa.array(rng.integers(90_000, 10_500_000, n)),
                  "discount": pa.array(rng.integers(0, 11, n)),
                  "quantity": pa.array(rng.integers(1, 51, n).astype(np.int32)),
                  "flag_a": flag(["A", "N", "R"], 3), "flag_b": flag(["F", "O"], 2)})
    buf = io.BytesIO()
    pq.write_table(t, buf, compression="none", row_group_size=1 << 20, write_statistics=True)
    data = buf.getvalue()
    out_bytes = n * (4 + 8 + 8 + 4 + 2 * (4 + 1))  # decoded columns: values + utf8 offsets and bytes

    def step():
        f = ParquetFile(data)
        tab = ParquetTable(ctx, f, list(range(f.num_row_groups)), list(range(f.num_columns)))
        return f, tab

    f, tab = step()
    exp = pq.ParquetFile(io.BytesIO(data)).read_row_group(3)
    ok = tab.to_arrow(3).equals(exp.to_batches()[0]) if exp.num_rows else True
    tab.close()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        f, tab = step()
        tab.close()
    wall = (time.perf_counter() - t0) / steps
    ctx.kernel_timing(True)
    f, tab = step()
    tab.close()
    kern = {}
    for name in ("parquet_index", "parquet_gather"):
        tot, cnt = ctx.kernel_time(name)
        kern[name] = {"total_ms": tot, "launches": cnt}
    ctx.kernel_timing(False)
    dev_ms = sum(k["total_ms"] for k in kern.values())
    t0 = time.perf_counter()
    for _ in range(steps):
        pq.read_table(io.BytesIO(data))
    pa_wall = (time.perf_counter() - t0) / steps
    return {"workload": f"Parquet scan into HBM: {n} rows x 6 columns (date, price, discount, quantity, flag A, "
                        "flag B), UNCOMPRESSED, dictionary encoding, 1,048,576-row row groups, file bytes in host "
                        "memory", "file_bytes": len(data), "row_groups": f.num_row_groups, "equal_to_pyarrow": bool(ok),
            "ms_per_scan": wall * 1e3, "rows_per_s": n / wall, "file_GBps": len(data) / wall / 1e9,
            "device_kernels": kern, "device_ms": dev_ms,
            "device_output_GBps": out_bytes / (dev_ms * 1e-3) / 1e9 if dev_ms else None,
            "pyarrow_read_table_ms": pa_wall * 1e3}


def bench_service(ctx, n_requests: int, rates, seed: int = 7, conc_threads: int = 8) -> dict:
    """Served request stream: Poisson arrivals, rows log-uniform in [4096, 16M];
    50% group-by (1M-row aggregation and fused filter + group-by shapes), 40%
    probes against 8 resident build tables of 1M..32M unique keys, each
    returning its (probe row, build row) pairs, 10% LIKE '%abc%' filters over
    a Utf8 column of 8..64 random lowercase bytes, each returning its bitmap.
    Inputs are device-resident (slices of generated pools); every request
    writes its outputs to a region of its own.  The same schedule runs as
    "served" (the operator service: continuous batching, persistent tile-queue
    kernels, graph replay), "plain" (one request per launch sequence, in
    arrival order, one thread and stream) and, at saturation, "plain_concurrent"
    (the one-request-per-launch operators from `conc_threads` threads, each with
    its own context and stream).  Latency = result on the host - arrival."""
    import threading

    import numpy as np

    from prism_executor import _lib as L
    from prism_executor import datagen as dg
    from prism_executor import device as D
    from prism_executor.service import Service, filter_mask

    lib = ctx.lib
    S0 = dg.SEED
    NP = 16_000_000 + 4096
    DOMAIN = 64_000_000
    I64, I32, UTF8 = L.TYPE_INT64, L.TYPE_INT32, L.TYPE_UTF8
    k0, v0 = D.empty(ctx, NP, np.int64), D.empty(ctx, NP, np.int64)
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, k0.ptr, NP, S0, 1, 0, 1000))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, v0.ptr, NP, S0, 2, 0, 1_000_000))
    date, price, disc = D.empty(ctx, NP, np.int32), D.empty(ctx, NP, np.int64), D.empty(ctx, NP, np.int64)
    fa, fao, fb, fbo = (D.empty(ctx, NP, np.uint8), D.empty(ctx, NP + 1, np.int32), D.empty(ctx, NP, np.uint8),
                        D.empty(ctx, NP + 1, np.int32))
    L.check(lib.prism_datagen_uniform_i32(ctx.handle, date.ptr, NP, S0, 20, 0, 2557))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, price.ptr, NP, S0, 21, 90_000, 10_410_000))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, disc.ptr, NP, S0, 22, 0, 11))
    L.check(lib.prism_datagen_flag_utf8(ctx.handle, fa.ptr, fao.ptr, NP, S0, 25, dg.FLAG_A, len(dg.FLAG_A)))
    L.check(lib.prism_datagen_flag_utf8(ctx.handle, fb.ptr, fbo.ptr, NP, S0, 26, dg.FLAG_B, len(dg.FLAG_B)))
    pk, pp = D.empty(ctx, NP, np.int64), D.empty(ctx, NP, np.int64)
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, pk.ptr, NP, S0, 40, 0, DOMAIN))
    L.check(lib.prism_datagen_bits_i64(ctx.handle, pp.ptr, NP, S0, 41))
    sb, so = D.empty(ctx, NP * 64, np.uint8), D.empty(ctx, NP + 1, np.int32)
    tb = ctypes.c_uint64()
    L.check(lib.prism_datagen_lower_utf8(ctx.handle, sb.ptr, NP * 64, so.ptr, NP, S0, 42, 8, 64, ctypes.byref(tb)))
    sizes = [1, 2, 3, 5, 8, 12, 20, 32]
    tables = []
    for i, m in enumerate(sizes):
        n = m * 1_000_000
        kk, pl = D.empty(ctx, n, np.int64), D.empty(ctx, n, np.int32)
        L.check(lib.prism_datagen_permutation_i64(ctx.handle, kk.ptr, n, S0, 70 + i))
        L.check(lib.prism_datagen_uniform_i32(ctx.handle, pl.ptr, n, S0, 80 + i, -(1 << 31), 1 << 32))
        tables.append((kk, pl))
    ctx.sync()

    GB0 = dict(column_types=[I64, I64], keys=[0], values=[[("LOAD_I", 1)]],
               aggs=[("SUM", 0), ("COUNT", 0), ("MIN", 0), ("MAX", 0)])
    GB2 = dict(column_types=[I32, I64, I64, UTF8, UTF8], keys=[3, 4], values=[REVENUE_EXACT],
               aggs=[("SUM", 0), ("COUNT", 0), ("AVG", 0)], filter_prog=FILTER_LE_2470)
    LIKE = [("LIKE", 0, 0)]

    def gb0_cols(o):
        return [D.column(I64, k0, offset=o), D.column(I64, v0, offset=o)]

    def gb2_cols(o):
        return [D.column(I32, date, offset=o), D.column(I64, price, offset=o), D.column(I64, disc, offset=o),
                D.column(UTF8, fa, fao, offset=o), D.column(UTF8, fb, fbo, offset=o)]

    def str_cols(o):
        return [D.column(UTF8, sb, so, offset=o)]

    rng = np.random.default_rng(seed)
    reqs = []
    pair_total = mask_total = 0
    for _ in range(n_requests):
        n = int(np.exp(rng.uniform(np.log(4096), np.log(16_000_000))))
        off = int(rng.integers(0, NP - n))
        u = rng.random()
        kind = "gb0" if u < 0.25 else "gb2" if u < 0.5 else "probe" if u < 0.9 else "filter"
        t = int(rng.integers(0, len(sizes)))
        # output region of this request: pairs (room for 1.25x the expected
        # matches; the result counts every match) or bitmap words
        if kind == "probe":
            cap = int(n * sizes[t] * 1e6 / DOMAIN * 1.25) + 4096
            reqs.append((kind, n, off, t, pair_total, cap))
            pair_total += cap
        elif kind == "filter":
            words = (n + 63) // 64
            reqs.append((kind, n, off, t, mask_total, words))
            mask_total += words
        else:
            reqs.append((kind, n, off, t, 0, 0))
    total_rows = sum(r[1] for r in reqs)
    gap = rng.exponential(1.0, n_requests)
    out_p, out_b = D.empty(ctx, max(pair_total, 1), np.uint32), D.empty(ctx, max(pair_total, 1), np.uint32)
    masks = D.empty(ctx, max(mask_total, 1), np.uint64)

    plain_tabs = [D.JoinTable(ctx, kk, payload=pl) for kk, pl in tables]  # resident, like the service's

    class Lane:
        """A plain caller: its own context (stream) and probe result buffer."""

        def __init__(self, c):
            self.ctx = c
            self.res = D.empty(c, ctypes.sizeof(L.ProbeResult), np.uint8)

        def run(self, req):
            kind, n, off, t, o, cap = req
            c = self.ctx
            if kind in ("gb0", "gb2"):
                g = D.GroupBy(c, expected_groups=1000 if kind == "gb0" else 8, **(GB0 if kind == "gb0" else GB2))
                g.update(gb0_cols(off) if kind == "gb0" else gb2_cols(off), n, 0)
                k = g.finish()
                out = g.read(k)
                g.close()
                return (k, tuple(out["sum128"][0]), tuple(out["first_row"].tolist()))
            if kind == "probe":
                tab = plain_tabs[t]
                L.check(lib.prism_join_probe_i64(c.handle, tab.handle, pk.ptr + 8 * off, n, 0, pp.ptr + 8 * off,
                                                 None, out_p.ptr + 4 * o, out_b.ptr + 4 * o, cap, self.res.ptr))
                c.sync()
                r = L.ProbeResult.from_buffer_copy(self.res.numpy().tobytes())
                return (int(r.matches), r.probe_sum, r.build_sum)
            sel = filter_mask(c, LIKE, [UTF8], str_cols(off), n, masks[o:o + cap], strings=[b"%abc%"])
            return (sel,)

    main_lane = Lane(ctx)

    def run_plain(arrivals):
        lat, res = [], []
        t0 = time.perf_counter()
        for req, a in zip(reqs, arrivals):
            now = time.perf_counter() - t0
            if now < a:
                time.sleep(a - now)
            res.append(main_lane.run(req))
            lat.append((time.perf_counter() - t0 - a) * 1e3)
        return lat, time.perf_counter() - t0, res

    lanes = [Lane(D.Context(0)) for _ in range(conc_threads)]

    def run_plain_concurrent(arrivals):
        lat = [0.0] * len(reqs)
        res = [None] * len(reqs)
        errors = []
        t0 = time.perf_counter()

        def worker(j):
            try:
                for i in range(j, len(reqs), conc_threads):
                    now = time.perf_counter() - t0
                    if now < arrivals[i]:
                        time.sleep(arrivals[i] - now)
                    res[i] = lanes[j].run(reqs[i])
                    lat[i] = (time.perf_counter() - t0 - arrivals[i]) * 1e3
            except Exception as e:  # reported below
                errors.append(repr(e))

        th = [threading.Thread(target=worker, args=(j,)) for j in range(conc_threads)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        if errors:
            raise RuntimeError(errors[0])
        return lat, time.perf_counter() - t0, res

    S = Service(ctx, hbm_budget=96 << 30, arena_bytes=1 << 30)
    for i, (kk, pl) in enumerate(tables):
        S.register_table(i, kk.numpy(), pl.numpy())
    S.register_groupby(0, max_groups=1000, **GB0)
    S.register_groupby(2, max_groups=8, **GB2)
    S.register_filter(4, LIKE, [UTF8], strings=[b"%abc%"])

    def run_served(arrivals):
        tickets = [0] * len(reqs)
        delay = [0.0] * len(reqs)
        t0 = time.perf_counter()

        def submitter():
            for i, ((kind, n, off, t, o, cap), a) in enumerate(zip(reqs, arrivals)):
                now = time.perf_counter() - t0
                if now < a:
                    time.sleep(a - now)
                delay[i] = (time.perf_counter() - t0 - a) * 1e3
                if kind == "gb0":
                    tickets[i] = S.submit_groupby(0, gb0_cols(off), n)
                elif kind == "gb2":
                    tickets[i] = S.submit_groupby(2, gb2_cols(off), n)
                elif kind == "probe":
                    tickets[i] = S.submit_probe(t, pk[off:off + n], pp[off:off + n], probe_idx=out_p[o:o + cap],
                                                build_idx=out_b[o:o + cap], capacity=cap)
                else:
                    tickets[i] = S.submit_filter(4, str_cols(off), n, masks[o:o + cap])

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        S.drain()
        wall = time.perf_counter() - t0
        lat, res = [], []
        for i, tk in enumerate(tickets):
            r = S.wait(tk)
            if r.status != 0:
                raise RuntimeError(f"served request {i} failed with status {r.status}")
            lat.append(delay[i] + r.latency_ms)
            kind = reqs[i][0]
            if kind in ("gb0", "gb2"):
                g = S.read_groups(tk, 0 if kind == "gb0" else 2, r.count)
                res.append((int(r.count), tuple(g["sum128"][0]), tuple(g["first_row"].tolist())))
            elif kind == "probe":
                res.append((int(r.count), r.probe_sum, r.build_sum))
            else:
                res.append((int(r.count),))
            S.release(tk)
        return lat, wall, res

    def summary(lat, wall, res=None):
        a = np.asarray(lat)
        return {"rows_per_s": total_rows / wall, "requests_per_s": len(reqs) / wall, "p50_ms": float(np.percentile(a, 50)),
                "p99_ms": float(np.percentile(a, 99)), "wall_s": wall}

    # warm-up: builds tables, captures graphs, tunes tiles, compiles the
    # generated group-by kernels of both forms
    _, _, served_res = run_served(np.zeros(len(reqs)))
    for kind, n in (("gb0", 4096), ("gb2", 4096)):
        for ln in [main_lane] + lanes:
            ln.run((kind, n, 0, 0, 0, 0))
    # parity: every served result equals the plain operator's on the same input
    _, _, plain_res = run_plain(np.zeros(len(reqs)))
    mismatches = sum(1 for a, b in zip(served_res, plain_res) if a != b)
    truncated = sum(1 for r, q in zip(served_res, reqs) if q[0] == "probe" and r[0] > q[5])
    # closed form of a sample of probes: table t's keys are a permutation of
    # 0..m-1, so key k matches iff k < m; its pairs in probe order
    hk = pk.numpy()
    checked = bad = 0
    for i in [i for i, q in enumerate(reqs) if q[0] == "probe"][:40]:
        kind, n, off, t, o, cap = reqs[i]
        hit = np.nonzero(hk[off:off + n] < sizes[t] * 1_000_000)[0]
        got = out_p[o:o + min(cap, len(hit))].numpy() if len(hit) else np.zeros(0, np.uint32)
        bad += int(served_res[i][0] != len(hit) or not np.array_equal(got, hit[:len(got)].astype(np.uint32)))
        checked += 1
    out = {"workload": f"{n_requests} requests, rows log-uniform [4096, 16M] (total {total_rows}), 50% group-by "
                       "(1M-row aggregation / fused filter + group-by shapes), 40% probes of 8 resident tables "
                       "(1M..32M unique keys) returning their (probe row, build row) pairs, 10% LIKE '%abc%' over "
                       "8..64-byte lowercase Utf8 returning the bitmap; inputs resident in HBM",
           "served_equals_plain": mismatches == 0, "mismatched_requests": mismatches,
           "probe_pairs_truncated": truncated, "probe_closed_form_checked": checked, "probe_closed_form_bad": bad,
           "plain_concurrent_threads": conc_threads, "rates": {}}
    for rate in rates:
        arr = np.cumsum(gap) / rate
        arr -= arr[0]
        S.reset_stats()
        served = summary(*run_served(arr))
        served["service"] = {k: v for k, v in S.stats().items() if k.startswith("mean_") or k.startswith("max_")}
        out["rates"][str(rate)] = {"served": served, "plain": summary(*run_plain(arr))}
        print(f"[bench] served stream at {rate} req/s done", file=sys.stderr, flush=True)
    # saturation: every request submitted at once; three runs of each form,
    # the run with the median served / plain ratio reported
    z = np.zeros(len(reqs))
    runs = []
    S.reset_stats()
    for _ in range(3):
        sv, pl, pc = summary(*run_served(z)), summary(*run_plain(z)), summary(*run_plain_concurrent(z))
        runs.append((sv["rows_per_s"] / pl["rows_per_s"], sv, pl, pc))
    runs.sort(key=lambda x: x[0])
    mid = runs[1]
    sat = {"served": mid[1], "plain": mid[2], "plain_concurrent": mid[3], "throughput_ratio": mid[0],
           "throughput_ratio_vs_concurrent": mid[1]["rows_per_s"] / mid[3]["rows_per_s"],
           "throughput_ratios_all_runs": [r[0] for r in runs]}
    st = S.stats()
    sat["service"] = {k: st[k] for k in ("batches", "mean_batch_requests", "mean_batch_rows", "mean_batch_gpu_ms",
                                         "max_batch_gpu_ms", "mean_batch_kernel_ms", "mean_batch_launch_ms",
                                         "mean_batch_detect_ms", "graph_launches") if k in st}
    out["saturation"] = sat
    out["service_stats"] = S.stats()
    S.close()
    for t in plain_tabs:
        t.close()
    for ln in lanes:
        ln.res.free()
        ln.ctx.close()
    main_lane.res.free()
    for a in [k0, v0, date, price, disc, fa, fao, fb, fbo, pk, pp, sb, so, out_p, out_b, masks] + \
            [x for tp in tables for x in tp]:
        a.free()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--build-rows", type=int, default=1_000_000)
    ap.add_argument("--probe-rows", type=int, default=64_000_000)
    ap.add_argument("--cpu-sample", type=int, default=16 * 1_048_576)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--groupby-rows", type=int, default=200_000_000)
    ap.add_argument("--no-groupby", action="store_true")
    ap.add_argument("--no-layouts", action="store_true", help="skip the join layout tuner sections")
    ap.add_argument("--no-join-agg", action="store_true")
    ap.add_argument("--join-agg-probe-rows", type=int, default=256_000_000)
    ap.add_argument("--no-service", action="store_true")
    ap.add_argument("--no-parquet", action="store_true")
    ap.add_argument("--parquet-rows", type=int, default=20_000_000)
    ap.add_argument("--svc-requests", type=int, default=10000)
    ap.add_argument("--svc-rates", type=str, default="200,500,1000")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="also run D2D copies of this many bytes")
    ap.add_argument("--skip-host-check", action="store_true",
                    help="skip the host recomputation of the 200M-row group-by sums (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); multi-GPU sharding is a design note, "
              "not a measured path", file=sys.stderr)
        return 2

    import numpy as np

    from prism_executor import _lib as L
    from prism_executor import datagen as dg
    from prism_executor import device as D

    nb, npr = args.build_rows, args.probe_rows
    t_start = time.perf_counter()
    ctx = D.Context(0)
    lib = ctx.lib
    bk = D.empty(ctx, nb, np.int64)
    bp = D.empty(ctx, nb, np.int32)
    pk = D.empty(ctx, npr, np.int64)
    pp = D.empty(ctx, npr, np.int64)
    L.check(lib.prism_datagen_permutation_i64(ctx.handle, bk.ptr, nb, dg.SEED, 10))
    L.check(lib.prism_datagen_uniform_i32(ctx.handle, bp.ptr, nb, dg.SEED, 11, -(1 << 31), 1 << 32))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, pk.ptr, npr, dg.SEED, 12, 0, 2 * nb))
    L.check(lib.prism_datagen_bits_i64(ctx.handle, pp.ptr, npr, dg.SEED, 13))
    op = D.empty(ctx, npr, np.uint32)
    ob = D.empty(ctx, npr, np.uint32)
    res = D.empty(ctx, 48, np.uint8)
    ctx.sync()

    def step(prev):
        if prev is not None:
            prev.close()
        t = D.JoinTable(ctx, bk, payload=bp)  # the build payload rides in the table entries
        t.probe_async(pk, res, pp, None, op, ob, capacity=npr)
        return t

    t = None
    for _ in range(args.warmup):
        t = step(t)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = step(t)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3

    r = L.ProbeResult.from_buffer_copy(res.numpy().tobytes())
    hpk = pk.numpy()
    sel = hpk < nb
    expected = int(sel.sum())
    hbk, hbp, hpp = bk.numpy(), bp.numpy(), pp.numpy()[sel]
    inv = np.empty(nb, np.int64)
    inv[hbk] = np.arange(nb)
    exp_build = int(hbp[inv[hpk[sel]]].astype(np.int64).sum())
    exp_probe = int((hpp >> 32).sum()) * (1 << 32) + int((hpp & 0xFFFFFFFF).sum())
    if r.error or r.matches != expected or r.build_sum != exp_build or r.probe_sum != exp_probe:
        print(f"probe result check failed: matches {r.matches} expected {expected} error {r.error}", file=sys.stderr)
        return 1

    # Whole-probe duration (events around K probes), then per-kernel
    # durations (events around each launch of the probe's kernels).
    ctx.timer_begin()
    for _ in range(args.steps):
        t.probe_async(pk, res, pp, None, op, ob, capacity=npr)
    probe_ms = ctx.timer_end() / args.steps
    ctx.kernel_timing(True)
    for _ in range(args.steps):
        t.probe_async(pk, res, pp, None, op, ob, capacity=npr)
    kernels = {}
    for name in ("join_probe", "join_probe_move"):
        tot, cnt = ctx.kernel_time(name)
        kernels[name] = {"avg_us": tot / max(cnt, 1) * 1e3, "launches": cnt}
    ctx.kernel_timing(False)
    # Build duration (host clock; the build synchronises internally).
    tb = time.perf_counter()
    for _ in range(5):
        t2 = D.JoinTable(ctx, bk, payload=bp)
        t2.close()
    build_ms = (time.perf_counter() - tb) / 5 * 1e3
    info = t.info
    t.close()

    alg_bytes = npr * 16 + r.matches * 8
    kern_ms = kernels["join_probe"]["avg_us"] * 1e-3
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    kernels["join_probe"]["alg_bytes"] = alg_bytes
    if kernels["join_probe_move"]["launches"]:  # the staged probe form moves the matches in a second kernel
        # staged build rows (4 B per match) and match ballots (512 B per 4096-row tile) read, pairs written
        move_bytes = r.matches * 12 + -(-npr // 4096) * 512
        kernels["join_probe_move"]["alg_bytes"] = move_bytes
        kernels["join_probe_move"]["GBps"] = move_bytes / (kernels["join_probe_move"]["avg_us"] * 1e-6) / 1e9
    else:
        del kernels["join_probe_move"]
    line = {
        "metric": METRIC,
        "value": npr / (ms_per_step * 1e-3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic, seeded on device: build keys = permutation of 0..1M-1, probe keys U[0,2M)",
        "config": {"workload": f"inner equi-join int64: build {nb} rows (unique keys, int32 payload) + probe {npr} "
                               f"rows (int64 payload) -> ordered row-index pairs + int128 payload sums",
                   "build_rows": nb, "probe_rows": npr, "table": "dense" if info.kind == L.TABLE_DENSE else "hashed"},
        "roofline": {"bound": "hbm", "kernel": "join_probe", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC.get("join_probe") if (nb, npr) == DEFAULT_SIZES[:2] else None},
        "roofline_whole_probe": {"bound": "hbm", "kernel": "join_probe + scan + join_probe_move (one probe call)",
                                 "achieved": alg_bytes / (probe_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": alg_bytes / (probe_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "kernels": kernels,
        "probe_ms": probe_ms,
        "probe_GBps": alg_bytes / (probe_ms * 1e-3) / 1e9,
        "probe_rows_per_s": npr / (probe_ms * 1e-3),
        "build_ms": build_ms,
        "matches": int(r.matches),
    }
    progress("join", t_start)
    if not args.no_layouts:
        line["join_layouts"] = bench_join_layouts(
            ctx, min(args.steps, 5), bk, bp, pk, pp, f"config above: build {nb} keys (permutation), probe {npr} rows")
        line["join_large_hashed"] = bench_large_hashed_join(ctx, min(args.steps, 5))
        progress("join layouts", t_start)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(bk.numpy(), pk[:args.cpu_sample].numpy())
        progress("join cpu baseline", t_start)
    for a in (op, ob, pk, pp):
        a.free()
    if not args.no_groupby:
        line["groupby_1m"] = bench_groupby_1m(ctx, args.steps, not args.no_cpu_baseline)
        progress("group-by 1M", t_start)
        line["groupby_filter"] = bench_groupby(ctx, args.groupby_rows, args.steps,
                                               0 if args.no_cpu_baseline else 16 * 1_048_576,
                                               check=not args.skip_host_check)
        progress("group-by", t_start)
    if not args.no_join_agg:
        line["join_aggregate"] = bench_join_agg(ctx, 16_000_000, args.join_agg_probe_rows, 8_000_000, 1_000_000,
                                                max(2, min(args.steps, 5)),
                                                0 if args.no_cpu_baseline else 20_000_000)
        progress("join-aggregate", t_start)
    if not args.no_parquet:
        line["parquet_scan"] = bench_parquet(ctx, args.parquet_rows, max(2, min(args.steps, 5)))
        progress("parquet scan", t_start)
    if not args.no_service:
        line["served_stream"] = bench_service(ctx, args.svc_requests,
                                              [int(x) for x in args.svc_rates.split(",") if x.strip()])
        progress("served stream", t_start)
    if args.calibrate_copy:
        src = D.empty(ctx, args.calibrate_copy, np.uint8)
        dst = D.empty(ctx, args.calibrate_copy, np.uint8)
        ctx.timer_begin()
        for _ in range(3):
            L.check(lib.prism_copy_device(ctx.handle, dst.ptr, src.ptr, args.calibrate_copy))
        cms = ctx.timer_end() / 3
        line["copy_calibration"] = {"bytes": args.calibrate_copy, "ms": cms,
                                    "GBps": 2 * args.calibrate_copy / (cms * 1e-3) / 1e9}
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
