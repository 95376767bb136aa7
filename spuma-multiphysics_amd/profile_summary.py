#!/usr/bin/env python3
"""Text summaries of rocprofv3 databases (rocpd SQLite output of ROCm 7).

  python3 spuma-multiphysics_amd/profile_summary.py trace DB
      per-kernel calls, total and average duration (the --stats view)
  python3 spuma-multiphysics_amd/profile_summary.py pmc DB [KERNEL_SUBSTRING]
      per-kernel mean of every collected counter per dispatch
  python3 spuma-multiphysics_amd/profile_summary.py dispatches DB KERNEL_SUBSTRING
      every dispatch of the matching kernels in launch order: grid, workgroup, duration
"""
import sqlite3
import sys
from collections import defaultdict


def trace(db):
    con = sqlite3.connect(db)
    # the top_kernels view reports durations in microseconds
    rows = con.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    print(f"{'kernel':72s} {'calls':>6s} {'total ms':>12s} {'avg us':>12s} {'%':>7s}")
    for name, calls, total, avg, pct in rows:
        print(f"{name[:72]:72s} {calls:6d} {total / 1e3:12.3f} {avg:12.3f} {pct:7.2f}")


def pmc(db, pattern=None):
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, name, duration, counter_name, counter_value from pmc_events").fetchall()
    per = defaultdict(lambda: defaultdict(float))  # (kernel, dispatch) -> counter -> summed value
    dur = {}
    for did, name, d, counter, value in rows:
        if pattern and pattern not in name:
            continue
        per[(name, did)][counter] += value
        dur[(name, did)] = d
    kernels = defaultdict(list)
    for (name, did), counters in per.items():
        kernels[name].append((did, dur[(name, did)], counters))
    for name, items in kernels.items():
        items.sort(key=lambda t: t[0])  # launch order
        names = sorted({c for _, _, cs in items for c in cs})
        print(f"{name[:100]}")
        print(f"  dispatches {len(items)}, mean duration {sum(d for _, d, _ in items) / len(items) / 1e3:.3f} us")
        for c in names:
            vals = [cs.get(c, 0.0) for _, _, cs in items]
            print(f"  {c:28s} mean {sum(vals) / len(vals):16.1f}   per dispatch {[round(v, 1) for v in vals[:6]]}"
                  + (f" ... last {[round(v, 1) for v in vals[-4:]]}" if len(vals) > 6 else ""))
        durations = [d for _, d, _ in items]
        if len(durations) > 6:
            print(f"  {'duration us':28s} first {[round(d / 1e3, 1) for d in durations[:6]]} ... last "
                  f"{[round(d / 1e3, 1) for d in durations[-4:]]}")


def dispatches(db, pattern):
    con = sqlite3.connect(db)
    rows = con.execute("select name, dispatch_id, grid_x, workgroup_x, duration from kernels order by start").fetchall()
    print(f"{'#':>4s} {'dispatch':>8s} {'grid':>10s} {'wg':>5s} {'us':>12s}  kernel")
    k = 0
    for name, did, grid, wg, d in rows:
        if pattern in name:
            print(f"{k:4d} {did:8d} {grid:10d} {wg:5d} {d / 1e3:12.3f}  {name.split('(')[0]}")
            k += 1


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        trace(sys.argv[2])
    elif sys.argv[1] == "dispatches":
        dispatches(sys.argv[2], sys.argv[3])
    else:
        pmc(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else None)
