"""Summarise rocprofv3 CSV output per kernel and launch shape.

rocprofv3's own --stats table aggregates by kernel name; the lattice
kernels run at every level with the same name, so this groups dispatches by
(kernel, grid size) and reports count / average / min / max duration, and,
for --pmc runs, the average counter value per dispatch.

    python -m fluidloom_amd.rocprof_summary --trace run_kernel_trace.csv \
        --pmc fetch_counter_collection.csv --pmc write_counter_collection.csv --match k_collide --match k_copy16
"""
from __future__ import annotations

import argparse
import csv
import re
import sys
from collections import defaultdict


def _col(header, *cands):
    for c in cands:
        for h in header:
            if h.lower() == c.lower():
                return h
    for c in cands:
        for h in header:
            if c.lower() in h.lower():
                return h
    return None


def _grid(row, header):
    g = _col(header, "Grid_Size")
    if g and row.get(g, "") not in ("", None):
        return int(float(row[g]))
    tot = 1
    for a in ("Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"):
        c = _col(header, a)
        if c and row.get(c):
            tot *= int(float(row[c]))
    return tot


def _short(name):
    name = re.sub(r"\s+", " ", name)
    return name if len(name) < 110 else name[:107] + "..."


def trace_table(path, matches):
    with open(path, newline="") as f:
        rd = csv.DictReader(f)
        hd = rd.fieldnames
        kn, t0, t1 = _col(hd, "Kernel_Name"), _col(hd, "Start_Timestamp"), _col(hd, "End_Timestamp")
        agg = defaultdict(list)
        for row in rd:
            name = row[kn]
            if matches and not any(m in name for m in matches):
                continue
            agg[(name, _grid(row, hd))].append(int(row[t1]) - int(row[t0]))
    out = ["kernel | grid (threads) | dispatches | avg us | min us | max us"]
    for (name, grid), ds in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"{_short(name)} | {grid} | {len(ds)} | {sum(ds) / len(ds) / 1e3:.2f} | "
                   f"{min(ds) / 1e3:.2f} | {max(ds) / 1e3:.2f}")
    return out


def pmc_table(path, matches):
    with open(path, newline="") as f:
        rd = csv.DictReader(f)
        hd = rd.fieldnames
        kn, cn, cv = _col(hd, "Kernel_Name"), _col(hd, "Counter_Name"), _col(hd, "Counter_Value")
        agg = defaultdict(list)
        for row in rd:
            name = row[kn]
            if matches and not any(m in name for m in matches):
                continue
            agg[(name, _grid(row, hd), row[cn])].append(float(row[cv]))
    out = ["kernel | grid (threads) | counter | dispatches | avg per dispatch"]
    for (name, grid, c), vs in sorted(agg.items()):
        out.append(f"{_short(name)} | {grid} | {c} | {len(vs)} | {sum(vs) / len(vs):.1f}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", action="append", default=[])
    ap.add_argument("--pmc", action="append", default=[])
    ap.add_argument("--match", action="append", default=[])
    a = ap.parse_args()
    for t in a.trace:
        print(f"## kernel trace: {t}")
        print("\n".join(trace_table(t, a.match)))
    for p in a.pmc:
        print(f"## counters: {p}")
        print("\n".join(pmc_table(p, a.match)))


if __name__ == "__main__":
    sys.exit(main())
