"""Summarise a rocprofv3 kernel trace of bench.py into per-kernel averages
over a range of steps of the main workload.

    python -m cudadsmc_amd.trace_summary <run_kernel_trace.csv> [--skip N] [--take K] [--grid-x G] [--grid-y Y]

The bench runs the main workload first (loading, warm-up, timed steps, then
the developed-flow and sampling blocks), then the ensemble and the c1/c2
lines. Step kernels of the main workload are picked by the tile kernel's
grid (Grid_Size_X = G threads, Grid_Size_Y = Y: 1 for one solver, the
member count for a batch); --skip drops the first N such steps and --take
keeps the next K.
"""
from __future__ import annotations

import argparse
import csv
from collections import defaultdict


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--skip", type=int, default=0)
    ap.add_argument("--take", type=int, default=20)
    ap.add_argument("--grid-x", type=int, default=262144)
    ap.add_argument("--grid-y", type=int, default=1)
    a = ap.parse_args()
    rows = list(csv.DictReader(open(a.trace)))
    for r in rows:  # template kernels are listed as "void name<...>(...)"
        if r["Kernel_Name"].startswith("void "):
            r["Kernel_Name"] = r["Kernel_Name"][5:]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    # step boundaries: each k_bscan_reduce launch with which=0 starts a step;
    # a step of the main workload is the run of kernels up to the next one,
    # identified by its tile kernel's grid
    steps, cur = [], []
    for r in rows:
        if r["Kernel_Name"].startswith("k_bscan_reduce") and cur and any(x["Kernel_Name"].startswith("k_tile") for x in cur):
            steps.append(cur)
            cur = []
        cur.append(r)
    if cur:
        steps.append(cur)
    main_steps = [s for s in steps if any(x["Kernel_Name"].startswith("k_tile") and int(x["Grid_Size_X"]) == a.grid_x
                                          and int(x["Grid_Size_Y"]) == a.grid_y for x in s)]
    sel = main_steps[a.skip:a.skip + a.take]
    tot = defaultdict(list)
    launches = defaultdict(int)
    for s in sel:
        per = defaultdict(float)
        for x in s:
            name = x["Kernel_Name"].split("(")[0]
            per[name] += (int(x["End_Timestamp"]) - int(x["Start_Timestamp"])) / 1e3
            launches[name] += 1
        for k, v in per.items():
            tot[k].append(v)
    print(f"main-workload steps in trace: {len(main_steps)}; summarised: steps {a.skip + 1}-{a.skip + len(sel)}")
    print("| kernel | launches per step | average us per launch | us per step |")
    print("|---|---|---|---|")
    total = 0.0
    for k, v in sorted(tot.items(), key=lambda kv: -sum(kv[1])):
        avg = sum(v) / len(sel)
        total += avg
        print(f"| `{k}` | {launches[k] / len(sel):.2f} | {sum(v) / launches[k]:.1f} | {avg:.1f} |")
    print(f"| sum | | | {total:.1f} |")


if __name__ == "__main__":
    main()
