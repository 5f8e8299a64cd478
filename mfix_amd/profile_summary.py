"""Summarise rocprofv3 CSV output directories (kernel trace stats and PMC
counter collections) into per-kernel averages.

    python -m mfix_amd.profile_summary DIR [DIR ...]

For every ``*kernel_stats.csv`` found under a DIR: calls, total and average
duration per kernel.  For every ``*kernel_trace.csv``: per kernel and grid
size, the launches that did work apart from those that returned at once (a
solve's launches after it converged, whose blocks only read the done flag):
a launch counts as working when it lasted at least a quarter of the longest
launch of the same kernel and grid.  For every ``*counter_collection.csv``:
the mean value per dispatch of each counter per kernel (FETCH_SIZE /
WRITE_SIZE are in KiB per dispatch).
"""
from __future__ import annotations

import csv
import sys
from collections import defaultdict
from pathlib import Path


def _col(row: dict, *keys: str):
    for k in keys:
        if k in row:
            return row[k]
    low = {c.lower(): c for c in row}
    for k in keys:
        if k.lower() in low:
            return row[low[k.lower()]]
    return None


def kernel_stats(path: Path) -> list[tuple]:
    out = []
    with path.open() as f:
        for row in csv.DictReader(f):
            name = _col(row, "Name", "KernelName", "Kernel_Name")
            calls = int(float(_col(row, "Calls") or 0))
            total = float(_col(row, "TotalDurationNs") or 0.0)
            avg = float(_col(row, "AverageNs") or (total / max(1, calls)))
            out.append((name, calls, total, avg))
    out.sort(key=lambda t: -t[2])
    return out


def trace_working(path: Path) -> list[tuple]:
    durs = defaultdict(list)
    with path.open() as f:
        for row in csv.DictReader(f):
            name = _col(row, "Kernel_Name", "KernelName", "Name")
            grid = _col(row, "Grid_Size_X", "Grid_Size", "Grid_SizeX") or "?"
            t0, t1 = _col(row, "Start_Timestamp"), _col(row, "End_Timestamp")
            durs[(name, grid)].append(int(t1) - int(t0))
    out = []
    for (name, grid), v in durs.items():
        cut = 0.25 * max(v)
        w = sorted(x for x in v if x >= cut)
        out.append((name, grid, len(v), len(w), sum(w), sum(w) / len(w), w[len(w) // 2]))
    out.sort(key=lambda t: -t[4])
    return out


def counters(path: Path) -> dict:
    acc = defaultdict(lambda: [0, 0.0])
    with path.open() as f:
        for row in csv.DictReader(f):
            name = _col(row, "Kernel_Name", "KernelName", "Name")
            cname = _col(row, "Counter_Name", "CounterName")
            val = float(_col(row, "Counter_Value", "CounterValue") or 0.0)
            a = acc[(name, cname)]
            a[0] += 1
            a[1] += val
    return {k: (v[0], v[1] / max(1, v[0])) for k, v in acc.items()}


def main(argv: list[str]) -> int:
    for d in argv:
        root = Path(d)
        for p in sorted(root.rglob("*kernel_stats.csv")):
            print(f"# kernel stats: {p}")
            print(f"{'calls':>8} {'total_ms':>10} {'avg_us':>9}  kernel")
            for name, calls, total, avg in kernel_stats(p)[:25]:
                print(f"{calls:8d} {total / 1e6:10.2f} {avg / 1e3:9.2f}  {name}")
        for p in sorted(root.rglob("*kernel_trace.csv")):
            print(f"# working launches (>= 1/4 of the longest of the same kernel and grid): {p}")
            print(f"{'grid':>10} {'calls':>7} {'working':>8} {'total_ms':>10} {'avg_us':>9} {'median_us':>10}  kernel")
            for name, grid, n, nw, tot, avg, med in trace_working(p)[:30]:
                print(f"{grid:>10} {n:7d} {nw:8d} {tot / 1e6:10.2f} {avg / 1e3:9.2f} {med / 1e3:10.2f}  {name}")
        for p in sorted(root.rglob("*counter_collection.csv")):
            print(f"# counters: {p}")
            print(f"{'dispatches':>10} {'mean':>14}  counter  kernel")
            rows = sorted(counters(p).items(), key=lambda kv: (-kv[1][0] * kv[1][1]))
            for (name, cname), (n, mean) in rows[:30]:
                print(f"{n:10d} {mean:14.1f}  {cname}  {name}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
