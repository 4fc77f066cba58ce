"""Summaries of rocprofv3 CSV output for the committed profile notes.

    python -m prism_executor.profsum stats <kernel_stats.csv>
    python -m prism_executor.profsum pmc <counter_collection.csv> [--copy-bytes N] [--copy-kernel NAME]

``stats`` prints rocprofv3's per-kernel totals (name, calls, average and
total duration).  ``pmc`` averages each counter per kernel name (KiB per
dispatch for FETCH_SIZE / WRITE_SIZE) and, given the byte count of the plain
device copy the run also performed, the counter's bytes-per-byte factor on
that copy, which turns the other kernels' readings into HBM bytes.
"""
import argparse
import csv
import sys
from collections import defaultdict


def short(name: str, width: int = 70) -> str:
    name = name.replace("prism::", "")
    return name if len(name) <= width else name[: width - 3] + "..."


def stats(path: str) -> None:
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    print(f"{'kernel':70s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}")
    for r in rows:
        avg_us = float(r["AverageNs"]) / 1e3
        tot_ms = float(r["TotalDurationNs"]) / 1e6
        print(f"{short(r['Name']):70s} {int(r['Calls']):6d} {avg_us:10.1f} {tot_ms:10.2f} {float(r['Percentage']):6.2f}")


def pmc(path: str, copy_bytes: int, copy_kernel: str = "k_copy16") -> None:
    acc = defaultdict(lambda: defaultdict(list))
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    factors = {}
    if copy_bytes:
        for name, ctrs in acc.items():
            if name.split("(")[0].split("::")[-1] == copy_kernel:
                for c, vals in ctrs.items():
                    kib = max(vals)
                    # a copy reads copy_bytes (FETCH) and writes copy_bytes (WRITE)
                    factors[c] = copy_bytes / (kib * 1024.0)
    print(f"{'kernel':70s} {'counter':>11s} {'n':>4s} {'avg_KiB':>14s} {'avg_MB(scaled)':>15s}")
    for name, ctrs in sorted(acc.items()):
        for c, vals in sorted(ctrs.items()):
            avg = sum(vals) / len(vals)
            scaled = avg * 1024.0 * factors.get(c, 1.0) / 1e6
            print(f"{short(name):70s} {c:>11s} {len(vals):4d} {avg:14.1f} {scaled:15.1f}")
    for c, fct in factors.items():
        print(f"copy calibration: {c} factor {fct:.4f} (copy of {copy_bytes} bytes)")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "pmc"])
    ap.add_argument("csv")
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--copy-kernel", default="k_copy16")
    a = ap.parse_args()
    if a.mode == "stats":
        stats(a.csv)
    else:
        pmc(a.csv, a.copy_bytes, a.copy_kernel)
    return 0


if __name__ == "__main__":
    sys.exit(main())
