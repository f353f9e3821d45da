"""
Helpers for hardware-counter runs under rocprofv3.

``python -m tad_dftd4_amd.profiling copy [MiB]`` runs a plain device copy of
a known size (default 1024 MiB, ten times) so that the FETCH_SIZE and
WRITE_SIZE counters can be scaled to bytes: bytes per dispatch of a kernel
= counter value x (known copy bytes / copy counter value).

``python -m tad_dftd4_amd.profiling summarize KERNEL_STATS_CSV [COUNTER_CSV ...]``
prints a short per-kernel table from rocprofv3 CSV output.
"""

from __future__ import annotations

import csv
import sys
from collections import defaultdict


def copy(mib: int = 1024, reps: int = 10) -> None:
    import torch

    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}", flush=True)


def _short(name: str) -> str:
    for pre in ("void ", "d4::(anonymous namespace)::", "(anonymous namespace)::", "d4::"):
        name = name.replace(pre, "")
    return name.split("(")[0].strip()


def summarize(stats_csv: str, *counter_csvs: str) -> None:
    with open(stats_csv, newline="") as f:
        rows = list(csv.DictReader(f))
    print(f"{'kernel':44s} {'calls':>6s} {'total_us':>11s} {'avg_us':>9s} {'pct':>6s}")
    for r in rows:
        tot = float(r["TotalDurationNs"]) / 1e3
        print(f"{_short(r['Name'])[:44]:44s} {int(r['Calls']):6d} {tot:11.1f} "
              f"{float(r['AverageNs']) / 1e3:9.1f} {float(r['Percentage']):6.2f}")
    for path in counter_csvs:
        acc: dict = defaultdict(lambda: defaultdict(list))
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                acc[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        print(f"\n# counters from {path.split('/')[-1]} (mean per dispatch)")
        for k, cs in sorted(acc.items()):
            vals = " ".join(f"{c}={sum(v) / len(v):.6g}" for c, v in sorted(cs.items()))
            print(f"{k[:44]:44s} {vals}")


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "copy"
    if what == "copy":
        copy(int(sys.argv[2]) if len(sys.argv) > 2 else 1024)
    elif what == "summarize":
        summarize(*sys.argv[2:])
    else:
        raise SystemExit(f"unknown command {what}")
