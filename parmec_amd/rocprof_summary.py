"""Summarise a rocprofv3 kernel trace (``--kernel-trace --output-format csv``)
per kernel and launch shape, optionally over the last N dispatches of each.

    python -m parmec_amd.rocprof_summary <prefix>_kernel_trace.csv [--last N] [--min-grid G]

Launches smaller than --min-grid threads (another, smaller workload of the
same run) are left out; --last N keeps the last N dispatches of each kernel
(the timed window of a benchmark follows its warm-up).
"""
import argparse
import csv
import re
from collections import defaultdict


def short(name):
    m = re.search(r"::(\w+)", name) or re.search(r"(\w+)", name)
    return m.group(1) if m else name


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--last", type=int, default=0)
    ap.add_argument("--min-grid", type=int, default=0)
    args = ap.parse_args()
    groups = defaultdict(list)
    for r in csv.DictReader(open(args.trace)):
        g = int(r["Grid_Size_X"])
        if g < args.min_grid:
            continue
        dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        groups[short(r["Kernel_Name"])].append((int(r["Start_Timestamp"]), dur, g))
    rows = []
    for k, v in groups.items():
        v.sort()
        if args.last:
            v = v[-args.last:]
        d = [x[1] for x in v]
        gs = [x[2] for x in v]
        rows.append((sum(d), k, min(gs), max(gs), len(d), sum(d) / len(d), min(d), max(d)))
    rows.sort(reverse=True)
    print("| kernel | grid (threads) | dispatches | avg us | min us | max us | total ms |")
    print("|---|---|---|---|---|---|---|")
    for tot, k, g0, g1, n, avg, mn, mx in rows:
        grid = str(g0) if g0 == g1 else f"{g0}-{g1}"
        print(f"| {k} | {grid} | {n} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} | {tot / 1e6:.2f} |")


if __name__ == "__main__":
    main()
