"""Summarise a rocprofv3 kernel trace (``--kernel-trace --output-format csv``)
per kernel and launch shape, optionally over the last N dispatches of each.

    python -m parmec_amd.rocprof_summary <prefix>_kernel_trace.csv [--last N] [--min-grid G]
    python -m parmec_amd.rocprof_summary <prefix>_counter_collection.csv --counters [--last N] [--min-grid G]

Launches smaller than --min-grid threads (another, smaller workload of the
same run) are left out; --last N keeps the last N dispatches of each kernel
(the timed window of a benchmark follows its warm-up).  With --counters the
file is a ``--pmc`` collection: the mean value per dispatch of every counter
for each kernel.
"""
import argparse
import csv
import re
from collections import defaultdict


def short(name):
    m = re.search(r"::(\w+)", name) or re.search(r"(\w+)", name)
    return m.group(1) if m else name


def grid_of(r):
    for k in ("Grid_Size_X", "Grid_Size"):
        if k in r and r[k] != "":
            return int(r[k])
    return 0


def counters(args):
    """Mean counter value per dispatch, per kernel (last N dispatches)."""
    per = defaultdict(dict)   # (kernel, dispatch) -> {counter: value}
    order = {}
    for r in csv.DictReader(open(args.trace)):
        if grid_of(r) < args.min_grid:
            continue
        key = (short(r["Kernel_Name"]), int(r["Dispatch_Id"]))
        per[key][r["Counter_Name"]] = per[key].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        order[key] = (grid_of(r), int(r["Dispatch_Id"]))
    kernels = defaultdict(list)
    for (k, d), vals in per.items():
        kernels[k].append((d, order[(k, d)][0], vals))
    names = sorted({c for v in per.values() for c in v})
    print("| kernel | grid (threads) | dispatches | " + " | ".join(f"{c} per dispatch" for c in names) + " |")
    print("|---|---|---|" + "---|" * len(names))
    for k, v in sorted(kernels.items()):
        v.sort()
        if args.last:
            v = v[-args.last:]
        gs = [x[1] for x in v]
        grid = str(gs[0]) if min(gs) == max(gs) else f"{min(gs)}-{max(gs)}"
        means = [sum(x[2].get(c, 0.0) for x in v) / len(v) for c in names]
        print(f"| {k} | {grid} | {len(v)} | " + " | ".join(f"{m:.6g}" for m in means) + " |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--last", type=int, default=0)
    ap.add_argument("--min-grid", type=int, default=0)
    ap.add_argument("--counters", action="store_true")
    args = ap.parse_args()
    if args.counters:
        counters(args)
        return
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
