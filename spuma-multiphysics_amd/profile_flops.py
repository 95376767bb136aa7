#!/usr/bin/env python3
"""FP32/FP64 VALU FLOPs per phase evaluation of the flash kernels, from a
rocprofv3 --pmc database of `profile_drivers.py uniform` (every lane of a wave
does the same work, so wave-level instruction counts x 64 are lane FLOPs):

  python3 spuma-multiphysics_amd/profile_flops.py DB FP64_PHASES FP32_PHASES

FP64_PHASES / FP32_PHASES are the phase evaluations the driver printed (both
flashes of the step); FLOPs count ADD, MUL and TRANS as 1 and FMA as 2."""
import sqlite3
import sys
from collections import defaultdict

PEAK = {"F64": 78.6e12, "F32": 157.3e12}


def kernels(db):
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, name, duration, counter_name, counter_value from pmc_events").fetchall()
    per = defaultdict(lambda: defaultdict(lambda: defaultdict(float)))
    dur = defaultdict(dict)
    for did, name, d, counter, value in rows:
        if "stepFlash" in name or "stepSearch" in name or "stepRefine" in name:
            key = name.split("<")[0].split("::")[-1]
            per[key][did][counter] += value
            dur[key][did] = d
    return per, dur


def main():
    db, p64, p32 = sys.argv[1], float(sys.argv[2]), float(sys.argv[3])
    per, dur = kernels(db)
    for name, dispatches in per.items():
        n = len(dispatches)
        mean = defaultdict(float)
        for counters in dispatches.values():
            for c, v in counters.items():
                mean[c] += v / n
        seconds = sum(dur[name].values()) / n / 1e9
        flop = {p: 64 * (mean[f"SQ_INSTS_VALU_ADD_{p}"] + mean[f"SQ_INSTS_VALU_MUL_{p}"] +
                         2 * mean[f"SQ_INSTS_VALU_FMA_{p}"] + mean[f"SQ_INSTS_VALU_TRANS_{p}"]) for p in PEAK}
        prec = "F32" if "Search" in name else "F64"
        phases = (p32 if prec == "F32" else p64) / 2  # per launch: two flashes per step
        print(f"{name}: {n} launches, {seconds * 1e3:.3f} ms each; {prec} FLOP per launch {flop[prec]:.4g}, "
              f"per phase evaluation {flop[prec] / phases:.0f}; {flop[prec] / seconds / 1e12:.2f} TFLOP/s = "
              f"{100 * flop[prec] / seconds / PEAK[prec]:.2f} % of the {PEAK[prec] / 1e12:.1f} TFLOP/s {prec} peak "
              f"(other precision: {flop['F64' if prec == 'F32' else 'F32']:.3g} FLOP)")


if __name__ == "__main__":
    main()
