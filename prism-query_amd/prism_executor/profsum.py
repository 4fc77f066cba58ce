"""Summaries of rocprofv3 CSV output for the committed profile notes.

    python -m prism_executor.profsum stats <kernel_stats.csv>
    python -m prism_executor.profsum trace <kernel_trace.csv> [--min-us A] [--max-us B]
    python -m prism_executor.profsum pmc <counter_collection.csv> [--copy-bytes N] [--copy-kernel NAME]
                                         [--min-us A] [--max-us B]

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


BYTE_COUNTERS = ("FETCH_SIZE", "WRITE_SIZE")  # KiB per dispatch; the others are event counts


def _keep(r, min_us: float, max_us: float) -> bool:
    us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    return min_us <= us <= max_us


def trace(path: str, min_us: float = 0.0, max_us: float = float("inf")) -> None:
    """Per kernel and grid size: dispatches and their mean / min / max duration
    (dispatches outside [min_us, max_us] left out)."""
    acc = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if _keep(r, min_us, max_us):
                us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
                acc[(r["Kernel_Name"], r["Grid_Size_X"])].append(us)
    print(f"{'kernel':70s} {'grid':>9s} {'n':>4s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s}")
    for (name, grid), v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        print(f"{short(name):70s} {grid:>9s} {len(v):4d} {sum(v) / len(v):10.1f} {min(v):10.1f} {max(v):10.1f}")


def pmc(path: str, copy_bytes: int, copy_kernel: str = "k_copy16", min_us: float = 0.0,
        max_us: float = float("inf")) -> None:
    acc = defaultdict(lambda: defaultdict(list))
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            if name.split("(")[0].split("::")[-1] != copy_kernel and not _keep(r, min_us, max_us):
                continue
            acc[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
    factors = {}
    if copy_bytes:
        for name, ctrs in acc.items():
            if name.split("(")[0].split("::")[-1] == copy_kernel:
                for c, vals in ctrs.items():
                    if c not in BYTE_COUNTERS:
                        continue
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
    derived(acc)


def derived(acc) -> None:
    """Ratios of counters collected in the same run, per kernel: L2 hit rate,
    mean L2 -> fabric read latency (cycles), LDS bank-conflict cycles per
    active LDS cycle, atomics per dispatch."""
    rows = []
    for name, c in sorted(acc.items()):
        avg = {k: sum(v) / len(v) for k, v in c.items()}
        out = []
        if "TCC_HIT" in avg and "TCC_MISS" in avg and avg["TCC_HIT"] + avg["TCC_MISS"] > 0:
            out.append(f"L2 hit {100 * avg['TCC_HIT'] / (avg['TCC_HIT'] + avg['TCC_MISS']):.1f}%")
        if avg.get("TCC_EA0_RDREQ"):
            if "TCC_EA0_RDREQ_LEVEL" in avg:
                out.append(f"EA read latency {avg['TCC_EA0_RDREQ_LEVEL'] / avg['TCC_EA0_RDREQ']:.0f} cyc")
            if "TCC_EA0_RDREQ_DRAM" in avg:
                out.append(f"EA reads to DRAM {100 * avg['TCC_EA0_RDREQ_DRAM'] / avg['TCC_EA0_RDREQ']:.1f}%")
        if avg.get("SQ_LDS_IDX_ACTIVE"):
            out.append(f"LDS bank-conflict cycles / active {avg.get('SQ_LDS_BANK_CONFLICT', 0) / avg['SQ_LDS_IDX_ACTIVE']:.3f}")
        if "TCC_ATOMIC" in avg:
            out.append(f"L2 atomics {avg['TCC_ATOMIC']:.3g}")
        if out:
            rows.append(f"{short(name):70s} " + "; ".join(out))
    if rows:
        print("derived:")
        for r in rows:
            print(r)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "trace", "pmc"])
    ap.add_argument("csv")
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--copy-kernel", default="k_copy16")
    ap.add_argument("--min-us", type=float, default=0.0, help="leave out shorter dispatches")
    ap.add_argument("--max-us", type=float, default=float("inf"), help="leave out longer dispatches")
    a = ap.parse_args()
    if a.mode == "stats":
        stats(a.csv)
    elif a.mode == "trace":
        trace(a.csv, a.min_us, a.max_us)
    else:
        pmc(a.csv, a.copy_bytes, a.copy_kernel, a.min_us, a.max_us)
    return 0


if __name__ == "__main__":
    sys.exit(main())
