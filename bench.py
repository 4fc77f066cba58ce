#!/usr/bin/env python3
"""Benchmark: inner equi-join on int64 keys on one MI355X.

Workload (one step): build a hash table over 1M build rows (unique keys = a
seeded permutation of 0..1M-1, int32 payload), then probe it with 64M rows
(keys U[0, 2M), ~50% match, int64 payload), writing the matched
(probe_row, build_row) index pairs in the reference's order plus the exact
int128 sums of both payloads over the matches.  Inputs are generated on the
device and stay resident in HBM; the timed region holds build + probe.

Prints one JSON line.  The roofline block covers the probe kernel (the
dominant one): algorithmic bytes = 16 B per probe row (key + payload read) +
8 B per matched pair (two u32 indices written), divided by the probe's
average duration measured with HIP events on the launch stream.
The CPU baseline times the oracle's restatement of the reference join on a
bounded sample (1M build rows, 8M probe rows) on one host core.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "prism-query_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "rows/s & %HBM peak (join probe, group-by); p99 latency batched vs plain"


def cpu_baseline(nb: int, sample: int) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as orc
    from prism_executor import datagen as dg

    d = dg.join_inputs(nb, sample)
    b = orc.Batch([orc.Col(orc.INT64, d["build_key"])])
    p = orc.Batch([orc.Col(orc.INT64, d["probe_key"])])
    t0 = time.perf_counter()
    pi, bi = orc.join_probe(b, [0], p, [0])
    dt = time.perf_counter() - t0
    assert len(pi) == int((d["probe_key"] < nb).sum())
    return {"value": sample / dt, "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"oracle C restatement of HashTable::build+probe, {nb} build rows + {sample} probe rows "
                      f"(first {sample} rows of the same seeded probe stream), one thread, {dt:.2f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--build-rows", type=int, default=1_000_000)
    ap.add_argument("--probe-rows", type=int, default=64_000_000)
    ap.add_argument("--cpu-sample", type=int, default=8_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); multi-GPU sharding is a design note, "
              "not a measured path", file=sys.stderr)
        return 2

    import torch

    from prism_executor import _lib as L
    from prism_executor import datagen as dg
    from prism_executor import device as D

    nb, npr = args.build_rows, args.probe_rows
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    ctx = D.Context(0)
    ctx.set_stream(stream.cuda_stream)
    lib = ctx.lib
    bk = torch.empty(nb, dtype=torch.int64, device="cuda")
    bp = torch.empty(nb, dtype=torch.int32, device="cuda")
    pk = torch.empty(npr, dtype=torch.int64, device="cuda")
    pp = torch.empty(npr, dtype=torch.int64, device="cuda")
    L.check(lib.prism_datagen_permutation_i64(ctx.handle, bk.data_ptr(), nb, dg.SEED, 10))
    L.check(lib.prism_datagen_uniform_i32(ctx.handle, bp.data_ptr(), nb, dg.SEED, 11, -(1 << 31), 1 << 32))
    L.check(lib.prism_datagen_uniform_i64(ctx.handle, pk.data_ptr(), npr, dg.SEED, 12, 0, 2 * nb))
    L.check(lib.prism_datagen_bits_i64(ctx.handle, pp.data_ptr(), npr, dg.SEED, 13))
    op = torch.empty(npr, dtype=torch.int32, device="cuda")
    ob = torch.empty(npr, dtype=torch.int32, device="cuda")
    res = torch.zeros(48, dtype=torch.uint8, device="cuda")
    ctx.sync()

    def step(prev):
        if prev is not None:
            prev.close()
        t = D.JoinTable(ctx, bk)
        t.probe_async(pk, res, pp, bp, op, ob, capacity=npr)
        return t

    t = None
    for _ in range(args.warmup):
        t = step(t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = step(t)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3

    r = L.ProbeResult.from_buffer_copy(res.cpu().numpy().tobytes())
    expected = int((pk < nb).sum())
    if r.error or r.matches != expected:
        print(f"probe result check failed: matches {r.matches} expected {expected} error {r.error}", file=sys.stderr)
        return 1

    # Probe kernel duration with events on the launch stream.
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for e0, e1 in evs:
        e0.record(stream)
        t.probe_async(pk, res, pp, bp, op, ob, capacity=npr)
        e1.record(stream)
    torch.cuda.synchronize()
    probe_ms = sum(e0.elapsed_time(e1) for e0, e1 in evs) / len(evs)
    # Build duration (host clock; the build synchronises internally).
    tb = time.perf_counter()
    for _ in range(5):
        t2 = D.JoinTable(ctx, bk)
        t2.close()
    build_ms = (time.perf_counter() - tb) / 5 * 1e3
    info = t.info
    t.close()

    alg_bytes = npr * 16 + r.matches * 8
    achieved = alg_bytes / (probe_ms * 1e-3) / 1e9
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "probe_ms": probe_ms,
        "probe_rows_per_s": npr / (probe_ms * 1e-3),
        "build_ms": build_ms,
        "matches": int(r.matches),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(nb, args.cpu_sample)
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
