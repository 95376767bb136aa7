#!/usr/bin/env python3
"""HLL + SSPRK2 + HEM-flash cell-steps/s on one MI355X.

Workload: pure N2O Peng-Robinson HEM in a 160^3 uniform Cartesian box (80 mm
cube, extrapolated boundaries, waveFactor 1.1, CFL 0.4), seeded state
T ~ U(180, 300) K, rho log-uniform 1-1000 kg/m^3, u ~ N(0, 20 m/s), internal
energy from the isochoric equilibrium at (rho, T).  One step = CFL, stage 0,
flash, CFL recheck, stage 1, flash, conservation totals, all resident on the
device (reactive_amd_step_run).  Inputs are resident in HBM before timing.

The timed flash is the mixed-precision one (FP32 search, FP64 refinement);
the plain FP64 path is timed on the same case for the comparison, and the
two are compared cell by cell after one step from the same states.

Prints one JSON line (the bench contract) with:
  roofline       the dominant kernel of the step (the mixed flash's FP32
                 candidate search, about two thirds of the step): VALU FLOPs
                 (phase evaluations x FLOPs per evaluation) / event-timed
                 launch duration vs the 157.3 TFLOP/s FP32 vector peak, and
                 its measured HBM traffic per launch
  roofline_stage fused flux sweep (both SSPRK2 stages): algorithmic bytes per
                 cell per stage / event-timed kernel duration vs 8 TB/s HBM
  roofline_flash all flash kernels: time per launch, phase evaluations
  fp64           the plain FP64 path on the same case
  mixed_vs_fp64  flips and relative T, p, c_f errors after one step
  cpu_baseline   the reference's own CPU code (oracle/_ref) on a bounded sample
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spuma-multiphysics_amd"))
from reactive_amd import abi, cases, lib as rlib, model  # noqa: E402

METRIC = "HLL+HEM-flash cell-steps/s at 160³: mixed vs FP64 GPU at ≤1e-10 rel T,p"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# the workloads: pure N2O (the metric's) and the binary N2O/N2 mixture
# (N2 mass fraction U(0, 0.3), T U(150, 320) K, rho log-uniform 0.5-1000)
CASES = {"n2o": {"species": ("N2O",), "cube": cases.pure_n2o_cube, "name": "n2o-pure"},
         "n2o-n2": {"species": ("N2O", "N2"), "cube": cases.binary_cube, "name": "n2o-n2-binary"}}


def prepare(n, seed, case="n2o"):
    """Seeded case + flash-consistent step-start states (one device flash)."""
    faces, volume, q, seeds, info = CASES[case]["cube"](n, seed=seed)
    ns = len(CASES[case]["species"])
    m = model.model_bytes(model.build_model(gas_species=CASES[case]["species"]))
    lib = rlib.load()
    err = C.create_string_buffer(1024)
    cells = len(q)
    hem = lib.reactive_gpu_hem_create_v1(m, len(m), cells, err, 1024)
    if not hem:
        raise RuntimeError(err.value.decode())
    e = cases.internal_energy(q, ns)
    states = seeds.copy()
    success = np.zeros(cells, np.int32)
    prof = abi.HemProfile.fresh()
    rc = lib.reactive_gpu_hem_run_v1(hem, rlib.ptr(np.ascontiguousarray(q[:, :ns])), rlib.ptr(e), cells,
                                     rlib.ptr(states), rlib.ptr(success), C.byref(prof), err, 1024)
    if rc:
        raise RuntimeError(err.value.decode())
    bad = np.where(success == 0)[0]
    if bad.size:  # a failed synthetic cell takes the state and conserved values of a converged neighbour
        good = np.where(success == 1)[0]
        if not good.size:
            raise RuntimeError("the device flash recovered none of the synthetic cells")
        src = good[np.searchsorted(good, bad).clip(0, good.size - 1)]
        q[bad] = q[src]
        states[bad] = states[src]
    info["prep_failures"] = int(bad.size)
    info["prep_flash_seconds"] = prof.kernelSeconds
    return faces, volume, q, states, m, hem, info


def host_cores():
    """The host's CPU description: nproc, the cores this process may use, the
    model name from lscpu."""
    out = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        import subprocess
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in text.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                out[key.strip()] = val.strip()
    except (OSError, ValueError):
        pass
    return out


class CpuPath:
    """The reference's own CPU code (oracle/_ref/libref.so, compiled from its
    sources) or, where that is absent, the oracle restatement: the flash over
    cells in `threads` workers (the reference's ThermoPool, one worker per
    core, reactivePool.inc:48-52) and the serial transport operators
    (backend 0)."""

    def __init__(self, threads):
        ref_path = os.path.join(ROOT, "oracle", "_ref", "libref.so")
        self.kind = "reference" if os.path.exists(ref_path) else "port"
        self.threads = threads
        V = C.c_void_p
        if self.kind == "reference":
            self.lib = C.CDLL(ref_path)
            self.lib.ref_flash_batch.argtypes = [V, C.c_size_t, V, V, V, C.c_int, V, V, V]
            self.lib.reactive_transport_create.restype = V
            self.lib.reactive_transport_create.argtypes = [C.c_int, V] + [V] * 6 + [C.c_char_p, C.c_size_t]
            self.lib.reactive_transport_stage.argtypes = [V] * 5 + [C.c_double, C.c_int, V]
            self.lib.reactive_transport_stable_step.argtypes = [V, V, V, C.c_double, C.c_double,
                                                                C.POINTER(C.c_double)]
            self.lib.reactive_transport_destroy.argtypes = [V]
        else:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            from helpers import oracle
            self.lib = oracle()

    def flash(self, m, q, e, seeds):
        """Flash every cell (q: cells x ns); returns (states, success, wall seconds)."""
        mb = C.create_string_buffer(m, len(m))
        q = np.ascontiguousarray(q)
        out = seeds.copy()
        succ = np.zeros(len(e), np.int32)
        chunks = np.array_split(np.arange(len(e)), self.threads * 8)

        def work(ch):
            if ch.size:
                a, b = ch[0], ch[-1] + 1
                if self.kind == "reference":
                    self.lib.ref_flash_batch(mb, b - a, rlib.ptr(q[a:b]), rlib.ptr(e[a:b]), rlib.ptr(seeds[a:b]), 0,
                                             rlib.ptr(out[a:b]), rlib.ptr(succ[a:b]), None)
                else:
                    self.lib.oracle_flash_batch(mb, b - a, rlib.ptr(q[a:b]), rlib.ptr(e[a:b]), rlib.ptr(out[a:b]),
                                                rlib.ptr(succ[a:b]), None, 1)
        t0 = time.perf_counter()
        with ThreadPoolExecutor(self.threads) as pool:
            list(pool.map(work, chunks))
        return out, succ, time.perf_counter() - t0

    def transport(self, faces, volume, ns):
        cfg = abi.Config(cells=len(volume), species=ns, variables=ns + 4, faces=len(faces), fixed=0, viscosity=0,
                         conductivity=0, diffusivity=0, waveFactor=1.1, maxBytes=1e12, mechanical=0)
        err = C.create_string_buffer(1024)
        if self.kind == "reference":
            h = self.lib.reactive_transport_create(0, C.byref(cfg), rlib.ptr(volume), rlib.ptr(faces), None, None,
                                                   None, None, err, 1024)
            if not h:
                raise RuntimeError(err.value.decode())
            lib = self.lib

            class T:
                def stage(self, q, st, dt, stage):
                    b = np.zeros(ns + 4)
                    assert lib.reactive_transport_stage(h, rlib.ptr(q), rlib.ptr(st), None, None, dt, stage,
                                                        rlib.ptr(b)) == 0

                def stable_step(self, q, st, cfl):
                    dt = C.c_double()
                    assert lib.reactive_transport_stable_step(h, rlib.ptr(q), rlib.ptr(st), cfl, 1.0,
                                                              C.byref(dt)) == 0
                    return dt.value

                def close(self):
                    lib.reactive_transport_destroy(h)
            return T()
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from helpers import OracleTransport
        ot = OracleTransport(cfg, volume, faces)

        class O:
            def stage(self, q, st, dt, stage):
                ot.stage(q, st, dt, stage)

            def stable_step(self, q, st, cfl):
                rc, dt = ot.stable_step(q, st, cfl, 1.0)
                assert rc == 0
                return dt

            def close(self):
                ot.close()
        return O()


def stratified_sample(states, size, rng):
    """Cells drawn per phase class (gas, two-phase, liquid by the step-start
    gas volume fraction) in proportion to the class's share of the box."""
    ag = states["alphaGas"]
    classes = {"gas": np.where(ag >= 1.0)[0], "two-phase": np.where((ag > 0) & (ag < 1))[0],
               "liquid": np.where(ag <= 0)[0]}
    n = len(states)
    picks, counts = [], {}
    for name, idx in classes.items():
        k = int(round(size * idx.size / n))
        k = min(k, idx.size)
        if k:
            picks.append(rng.choice(idx, size=k, replace=False))
        counts[name] = k
    return np.sort(np.concatenate(picks)), counts


def cpu_baseline(m, q_end, seeds_start, threads, sample, seed):
    """Reference CPU path on a bounded sample of the same workload: its flash
    on `sample` cells of a real step's second flash, stratified by phase class,
    in `threads` workers; its serial transport on a 48^3 box; per-cell costs
    scaled to cell-steps/s.  Also one whole step of the 40^3 pure-N2O box
    (CFL, stage 0, flash, stage 1, flash), timed end to end."""
    cpu = CpuPath(threads)
    rng = np.random.default_rng(seed)
    idx, counts = stratified_sample(seeds_start, min(sample, len(q_end)), rng)
    e = cases.internal_energy(q_end[idx], 1)
    _, succ, flash_wall = cpu.flash(m, q_end[idx, :1], e, seeds_start[idx].copy())
    # transport: serial operators on a 48^3 box, two stages
    tn = 48
    faces, volume, q0, seeds0, _ = cases.pure_n2o_cube(tn, seed=seed)
    tst = np.zeros(len(q0), dtype=abi.TSTATE_DTYPE)
    tst["p"] = 1e5; tst["T"] = 250.0; tst["rho"] = q0[:, 0]; tst["cv"] = 800.0; tst["sound"] = 200.0
    tst["gasMass"] = q0[:, 0]
    tr = cpu.transport(faces, volume, 1)
    qq = q0.copy()
    t0 = time.perf_counter()
    for stage in (0, 1):
        tr.stage(qq, tst, 1e-8, stage)
    transport_wall = time.perf_counter() - t0
    tr.close()
    flash_per_cell = flash_wall / len(idx)            # wall seconds per cell-flash with `threads` workers
    transport_per_cell = transport_wall / len(q0)     # serial seconds per cell for both stages
    per_cell_step = 2 * flash_per_cell + transport_per_cell
    # one full step of the 40^3 box
    n0 = 40
    faces, volume, q, seeds, _ = cases.pure_n2o_cube(n0, seed=seed)
    states, ok, _ = cpu.flash(m, q[:, :1], cases.internal_energy(q, 1), seeds)
    states[ok == 0] = states[np.where(ok == 1)[0][0]]
    tr = cpu.transport(faces, volume, 1)
    t0 = time.perf_counter()
    ts = cases.transport_states(states)
    dt = tr.stable_step(q, ts, 0.4)
    q0 = q.copy()
    tr.stage(q, ts, dt, 0)
    s1, ok1, _ = cpu.flash(m, q[:, :1], cases.internal_energy(q, 1), states)
    s1 = np.where(ok1 == 1, s1, states)
    tr.stage(q, cases.transport_states(s1), dt, 1)
    s2, ok2, _ = cpu.flash(m, q[:, :1], cases.internal_energy(q, 1), states)
    step_wall = time.perf_counter() - t0
    tr.close()
    del q0
    return {"value": 1.0 / per_cell_step, "unit": "cell-steps/s", "cores": threads, "kind": cpu.kind,
            "host": host_cores(),
            "sample": (f"{len(idx)} cells of a real 160^3 step's second flash, stratified by phase class "
                       f"({', '.join(f'{k} {v}' for k, v in counts.items())}; flash x2 per step, {threads} "
                       f"worker threads, {flash_wall:.2f} s wall) + serial transport, both SSPRK2 stages on a "
                       f"{tn}^3 box ({transport_wall:.2f} s); per-cell costs scaled; {int(succ.sum())}/{len(idx)} "
                       f"sample flashes converged"),
            "config0_step": {"workload": "cube40-n2o-pure, one whole step (CFL, stage 0, flash, stage 1, flash)",
                             "cells": len(q), "seconds": step_wall, "value": len(q) / step_wall,
                             "unit": "cell-steps/s", "flash_failures": [int((ok1 == 0).sum()), int((ok2 == 0).sum())]}}


def run_steps(lib, step, q, states, cfl, steps, mode, warmup, profile_steps):
    """Loads (q, states), runs `warmup` untimed steps (at least one, on the
    stream: it autotunes the launch shapes and records the per-cell costs that
    order the timed flashes; the others replay the captured graph), then times
    `steps` steps replayed from the captured HIP graph (REACTIVE_AMD_STEP_GRAPH:
    the metric's wall clock), then runs `profile_steps` steps on the stream for
    the per-kernel event times the rooflines use.  Returns (timed report, wall
    seconds, kernel-time report, q and states after the first warmup step)."""
    def check(rc):
        if rc:
            raise RuntimeError(lib.reactive_amd_step_error(step).decode())
    check(lib.reactive_amd_step_load(step, rlib.ptr(q), rlib.ptr(states)))
    rep = abi.StepReport.fresh()
    check(lib.reactive_amd_step_run(step, cfl, 1.0, 1, mode | AUTOTUNE, C.byref(rep)))
    q_one = np.zeros_like(q)
    s_one = np.zeros(len(q), dtype=abi.STATE_DTYPE)
    check(lib.reactive_amd_step_read(step, rlib.ptr(q_one), rlib.ptr(s_one)))
    # the first graph call captures (a graph per buffer parity): keep both captures out of the timing
    check(lib.reactive_amd_step_run(step, cfl, 1.0, max(2, warmup - 1), mode | GRAPH, C.byref(rep)))
    rep = abi.StepReport.fresh()
    t0 = time.perf_counter()
    check(lib.reactive_amd_step_run(step, cfl, 1.0, steps, mode | GRAPH, C.byref(rep)))
    wall = time.perf_counter() - t0
    kernels = abi.StepReport.fresh()
    check(lib.reactive_amd_step_run(step, cfl, 1.0, profile_steps, mode, C.byref(kernels)))
    return rep, wall, kernels, q_one, s_one


def stage_roofline(rep, cells, steps, n, nv=5):
    bytes0, bytes1 = 8 * (2 * nv + 2), 8 * (3 * nv + 2)  # algorithmic bytes per cell, stage 0 / stage 1
    t_stage = (rep.stageSeconds[0] + rep.stageSeconds[1]) / steps
    achieved = (bytes0 + bytes1) * cells / t_stage / 1e9
    return {"bound": "hbm", "kernel": "batchTiledStageKernel (both SSPRK2 stages)", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": STAGE_TRAFFIC_160 if n == 160 and nv == 5 else None,
            "bytes_per_cell": [bytes0, bytes1],
            "avg_launch_ms": [1e3 * rep.stageSeconds[0] / steps, 1e3 * rep.stageSeconds[1] / steps]}


# HBM bytes per launch of the FP32 search at 160^3 pure N2O (the bench's
# seeded box, mixed step, tuned launch shapes), from rocprofv3 FETCH_SIZE and
# WRITE_SIZE scaled as below: profiles/stepSearch160_traffic.txt
SEARCH_TRAFFIC_160 = 135.97e9
# The source commit the profiles behind SEARCH_TRAFFIC_160, STAGE_TRAFFIC_160
# and FLOP_PER_PHASE were measured on (reported with the rooflines).
PROFILED_COMMIT = "4d905b9"

# HBM bytes per launch of the fused stage kernel at 160^3 pure N2O, stage 0 and
# stage 1, from rocprofv3 FETCH_SIZE and WRITE_SIZE (each in its own run,
# scaled by a plain copy kernel's counts): profiles/cartesianStage160_traffic.txt
STAGE_TRAFFIC_160 = [473.1e6, 708.0e6]
TIGHT = {"vtol": 1e-12, "etol": 1e-12, "mutol": 1e-10}
AUTOTUNE = 0x200  # REACTIVE_AMD_STEP_AUTOTUNE
GRAPH = 0x100  # REACTIVE_AMD_STEP_GRAPH
# VALU FLOPs per phase evaluation (ADD, MUL, TRANS = 1, FMA = 2), measured with
# rocprofv3 on divergence-free uniform cells, mean of a two-phase cell and a
# gas cell: profiles/flash_valu_uniform32.txt.  The flash roofline's achieved
# rate is phase evaluations x this figure / event-timed kernel time.
FLOP_PER_PHASE = {"fp64_flash": (1619 + 2115) / 2, "fp32_search": (370 + 523) / 2, "fp64_refine": (2253 + 1596) / 2}
FP32_PEAK_TFLOPS = 157.3


def converged_roots(lib, q, seeds, case="n2o"):
    """The FP64 flash of (q, seeds) rerun on the device with tolerances 1e-3
    as tight: the converged root each flash's result is compared with."""
    ns = len(CASES[case]["species"])
    tight = model.model_bytes(model.build_model(limits=TIGHT, gas_species=CASES[case]["species"]))
    err = C.create_string_buffer(1024)
    cells = len(q)
    hem = lib.reactive_gpu_hem_create_v1(tight, len(tight), cells, err, 1024)
    if not hem:
        raise RuntimeError(err.value.decode())
    out = seeds.copy()
    ok = np.zeros(cells, np.int32)
    prof = abi.HemProfile.fresh()
    rc = lib.reactive_gpu_hem_run_v1(hem, rlib.ptr(np.ascontiguousarray(q[:, :ns])),
                                     rlib.ptr(cases.internal_energy(q, ns)), cells, rlib.ptr(out), rlib.ptr(ok),
                                     C.byref(prof), err, 1024)
    lib.reactive_gpu_hem_destroy_v1(hem)
    if rc:
        raise RuntimeError(err.value.decode())
    return out, ok == 1


def distance(s, states0, root, ok):
    cells = len(s)
    fail = (s.view(np.uint8).reshape(cells, -1) == states0.view(np.uint8).reshape(cells, -1)).all(1)
    use = ~fail & ok & (s["activeLiquids"] == root["activeLiquids"])
    out = {"compared_cells": int(use.sum()),
           "other_phase_set_than_root": int((~fail & ok & (s["activeLiquids"] != root["activeLiquids"])).sum())}
    for f in ("T", "p"):
        r = np.abs(s[f][use] - root[f][use]) / np.abs(root[f][use])
        out[f] = {"max": float(r.max()), "p99": float(np.percentile(r, 99)),
                  "share_le_1e-10": float((r <= 1e-10).mean()), "cells_over_1e-10": int((r > 1e-10).sum())}
    return out


def flash_roofline(rep, rep64, cells, steps, mixed):
    """VALU rate of the flash kernels against the vector peaks (see FLOP_PER_PHASE)."""
    flashes = 2 * steps
    out = {"bound": "valu", "unit": "TFLOP/s", "flop_source": "profiles/flash_valu_uniform32.txt",
           "profiled_commit": PROFILED_COMMIT}
    t64 = (rep64.flashSeconds[0] + rep64.flashSeconds[1]) / flashes
    a64 = rep64.phaseEvaluations / flashes * FLOP_PER_PHASE["fp64_flash"] / t64 / 1e12
    out["fp64_flash"] = {"kernel": "stepFlashKernel", "avg_launch_ms": 1e3 * t64, "achieved": a64,
                         "peak": FP64_PEAK_TFLOPS, "frac": a64 / FP64_PEAK_TFLOPS,
                         "phase_evaluations_per_cell": rep64.phaseEvaluations / (cells * flashes)}
    if mixed:
        t = (rep.flashSeconds[0] + rep.flashSeconds[1]) / flashes
        ts = rep.searchSeconds / flashes
        tr = t - ts
        a32 = rep.singlePhaseEvaluations / flashes * FLOP_PER_PHASE["fp32_search"] / ts / 1e12
        ar = rep.phaseEvaluations / flashes * FLOP_PER_PHASE["fp64_refine"] / tr / 1e12
        out["avg_launch_ms"] = 1e3 * t
        out["fp32_search"] = {"kernel": "stepSearchKernel", "avg_launch_ms": 1e3 * ts, "achieved": a32,
                              "peak": FP32_PEAK_TFLOPS, "frac": a32 / FP32_PEAK_TFLOPS,
                              "phase_evaluations_per_cell": rep.singlePhaseEvaluations / (cells * flashes)}
        out["fp64_refine"] = {"kernel": "stepRefineKernel (+ stepRefineQueueKernel)", "avg_launch_ms": 1e3 * tr,
                              "achieved": ar, "peak": FP64_PEAK_TFLOPS, "frac": ar / FP64_PEAK_TFLOPS,
                              "phase_evaluations_per_cell": rep.phaseEvaluations / (cells * flashes),
                              "queued_cells_per_flash": rep.queuedCells / flashes}
        out["share_of_step"] = (rep.flashSeconds[0] + rep.flashSeconds[1]) / rep.seconds
    else:
        out["avg_launch_ms"] = 1e3 * t64
        out["share_of_step"] = (rep64.flashSeconds[0] + rep64.flashSeconds[1]) / rep64.seconds
    return out


def search_roofline(flash, rep, steps, n, case="n2o"):
    """The dominant kernel: the FP32 candidate search of the mixed flash."""
    s = flash["fp32_search"]
    return {"bound": "valu", "kernel": "stepSearchKernel (FP32 candidate search of the mixed flash)",
            "achieved": s["achieved"], "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": s["frac"],
            "traffic": SEARCH_TRAFFIC_160 if n == 160 and case == "n2o" else None,
            "traffic_source": "profiles/stepSearch160_traffic.txt", "profiled_commit": PROFILED_COMMIT,
            "avg_launch_ms": s["avg_launch_ms"],
            "flop_per_phase_evaluation": FLOP_PER_PHASE["fp32_search"],
            "phase_evaluations_per_cell": s["phase_evaluations_per_cell"],
            "share_of_step": rep.searchSeconds / rep.seconds}


def census(states0, s64, smx):
    """Mixed vs FP64 after one step from the same states: cells that failed
    in one flash only or chose another phase set (flips), and the relative
    T, p error over the rest."""
    cells = len(s64)
    raw0 = states0.view(np.uint8).reshape(cells, -1)
    fail64 = (s64.view(np.uint8).reshape(cells, -1) == raw0).all(1)
    failmx = (smx.view(np.uint8).reshape(cells, -1) == raw0).all(1)
    both = ~fail64 & ~failmx
    flips = int((fail64 != failmx).sum() + (smx["activeLiquids"][both] != s64["activeLiquids"][both]).sum())
    same = both & (smx["activeLiquids"] == s64["activeLiquids"])
    out = {"flips": flips, "compared_cells": int(same.sum())}
    for f in ("T", "p", "soundFrozen"):
        r = np.abs(smx[f][same] - s64[f][same]) / np.abs(s64[f][same])
        out[f] = {"max": float(r.max()), "p99": float(np.percentile(r, 99)),
                  "share_le_1e-10": float((r <= 1e-10).mean())}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1, help="untimed steps (at least one is always run)")
    ap.add_argument("--n", type=int, default=160, help="cells per edge")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--flash", choices=["mixed", "fp64"], default="mixed",
                    help="timed flash arithmetic; mixed also times the FP64 path for the comparison")
    ap.add_argument("--profile-steps", type=int, default=3,
                    help="stream steps after the timed ones, for the per-kernel event times")
    ap.add_argument("--cpu-sample", type=int, default=65536)
    # the reference pool runs one worker per core up to 64 (reactivePool.inc:50); a one-GPU box's
    # CPU share is 16 cores
    ap.add_argument("--cpu-threads", type=int, default=min(16, 64, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--case", choices=sorted(CASES), default="n2o",
                    help="n2o: the metric's pure-N2O box; n2o-n2: the binary mixture box (no CPU baseline)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported (the multi-GPU design is "
              "documented, not built)", file=sys.stderr)
        sys.exit(2)
    lib = rlib.load()
    species = CASES[a.case]["species"]
    ns = len(species)
    log(f"bench: building {a.n}^3 {CASES[a.case]['name']} case (seed {a.seed})")
    faces, volume, q, states, m, hem, info = prepare(a.n, a.seed, a.case)
    cells = len(q)
    cfg = abi.Config(cells=cells, species=ns, variables=ns + 4, faces=len(faces), fixed=0, viscosity=0,
                     conductivity=0,
                     diffusivity=0, waveFactor=1.1, maxBytes=200e9, mechanical=0)
    err = C.create_string_buffer(1024)
    tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(volume), rlib.ptr(faces), None, None, None, None,
                                       err, 1024)
    if not tr:
        raise RuntimeError(err.value.decode())
    step = lib.reactive_amd_step_create(tr, hem, err, 1024)
    if not step:
        raise RuntimeError(err.value.decode())
    prof_steps = max(1, min(a.steps, a.profile_steps))
    log(f"bench: FP64 flash, {a.warmup} warmup + {a.steps} timed (graph) + {prof_steps} kernel-timed step(s)")
    rep64, wall64, kern64, q_one, s64 = run_steps(lib, step, q, states, a.cfl, a.steps, 0, a.warmup, prof_steps)
    q64_one = q_one
    log(f"bench: FP64 {wall64:.3f} s for {a.steps} steps (device {rep64.seconds:.3f} s)")
    value64 = cells * a.steps / wall64
    t_flash64 = (kern64.flashSeconds[0] + kern64.flashSeconds[1]) / (2 * prof_steps)
    fp64 = {"value": value64, "ms_per_step": 1e3 * wall64 / a.steps, "flash_ms": 1e3 * t_flash64,
            "phase_evaluations_per_cell_step": rep64.phaseEvaluations / (cells * a.steps),
            "flash_failures": [int(rep64.flashFailures[0]), int(rep64.flashFailures[1])],
            "flash_failure_share": [int(rep64.flashFailures[i]) / (cells * a.steps) for i in (0, 1)]}
    if a.flash == "mixed":
        log(f"bench: mixed flash, {a.warmup} warmup + {a.steps} timed (graph) + {prof_steps} kernel-timed step(s)")
        rep, wall, kern, qmx_one, smx = run_steps(lib, step, q, states, a.cfl, a.steps, 1, a.warmup, prof_steps)
        log(f"bench: mixed {wall:.3f} s for {a.steps} steps (device {rep.seconds:.3f} s)")
        accuracy = census(states, s64, smx)
        log("bench: converged roots of both steps' second flash")
        root, ok = converged_roots(lib, qmx_one, states, a.case)
        accuracy["mixed_vs_converged_root"] = distance(smx, states, root, ok)
        root, ok = converged_roots(lib, q64_one, states, a.case)
        accuracy["fp64_vs_converged_root"] = distance(s64, states, root, ok)
    else:
        rep, wall, kern, accuracy = rep64, wall64, kern64, None
    value = cells * a.steps / wall
    mixed = a.flash == "mixed"
    flash = flash_roofline(kern, kern64, cells, prof_steps, mixed)
    result = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32+fp64" if mixed else "f64", "data": "synthetic",
        "config": {"workload": f"cube{a.n}-{CASES[a.case]['name']}", "cells": cells, "faces": len(faces),
                   "cfl": a.cfl, "species": list(species), "liquids": ["N2O"],
                   "flash": ("mixed: FP32 candidate search, FP64 refinement, stability and selection" if mixed
                             else "fp64 (plain, reference arithmetic)"),
                   "seed": a.seed, "two_phase_fraction": (round(info["two_phase_fraction"], 4)
                                                          if "two_phase_fraction" in info else None),
                   "timed_path": "steps replayed from captured HIP graphs (REACTIVE_AMD_STEP_GRAPH)",
                   "kernel_times": f"HIP events on the step stream over {prof_steps} further step(s)"},
        "launch_shapes": {"stage_tile_depth": int(kern.tileDepth), "flash_block": int(kern.flashBlock),
                          "state_bins_per_axis": int(kern.stateBins), "state_cost_levels": int(kern.reserved),
                          "fp64_flash_block": int(kern64.flashBlock), "fp64_state_bins_per_axis": int(kern64.stateBins),
                          "fp64_state_cost_levels": int(kern64.reserved), "source": "autotuned in the untimed steps"},
        "roofline": (search_roofline(flash, kern, prof_steps, a.n, a.case) if mixed
                     else dict(flash["fp64_flash"], bound="valu", unit="TFLOP/s")),
        "roofline_stage": stage_roofline(kern, cells, prof_steps, a.n, ns + 4),
        "roofline_flash": flash,
        "fp64": fp64,
        "speedup_vs_fp64": value / value64,
        "mixed_vs_fp64": accuracy,
        "flash_failures": [int(rep.flashFailures[0]), int(rep.flashFailures[1])],
        # cells whose flash failed (they keep their step-start state), per flash and step; the
        # reference driver's attempt loop (ReactiveFoam.C:1487-1494) would reject such a step and
        # retry it at half the dt: these steps are counted as the resident step runs them (no retry)
        "flash_failure_share": [int(rep.flashFailures[i]) / (cells * a.steps) for i in (0, 1)],
        "cfl_violations": int(rep.cflViolations), "dt": rep.dt,
    }
    if not a.no_cpu and a.case == "n2o":
        log(f"bench: cpu baseline ({a.cpu_sample} cells, {a.cpu_threads} threads)")
        result["cpu_baseline"] = cpu_baseline(m, q_one, states, a.cpu_threads, a.cpu_sample, a.seed)
    lib.reactive_amd_step_destroy(step)
    lib.reactive_transport_destroy(tr)
    lib.reactive_gpu_hem_destroy_v1(hem)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
