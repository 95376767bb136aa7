#!/usr/bin/env python3
"""Small drivers for rocprofv3 passes (each runs in seconds):

  stage N REPS   fused Cartesian SSPRK2 stages (batchTiledStageKernel) on an N^3
                 pure-N2O box through reactive_transport_stage, REPS pairs
  flash N        one HEM flash batch (hemFlashKernel) of an N^3 box through
                 reactive_gpu_hem_run_v1
  cube N MODE STEPS
                 resident SSPRK2 steps (MODE 0 FP64, 1 mixed) on the seeded
                 pure-N2O N^3 box of the benchmark
  bench N MODE STEPS
                 the benchmark's own flow (bench.prepare + bench.run_steps) on its
                 seeded pure-N2O N^3 box: an autotuning step on the stream, two
                 graph-replayed warm-up steps, STEPS timed graph-replayed steps
                 and one step on the stream (MODE 0 FP64, 1 mixed), so the
                 kernels run with the launch shapes the benchmark times
  copy MIB       a plain device copy kernel of MIB MiB (torch elementwise copy),
                 to calibrate FETCH_SIZE and WRITE_SIZE against a known byte count
  uniform N T RHO MODE STEPS
                 resident SSPRK2 steps (MODE 0 FP64, 1 mixed) on an N^3 box whose
                 cells all hold the same state (T, RHO, at rest): every lane of a
                 flash wave does the same work, so VALU counters divided by the
                 phase evaluations give the operations per phase evaluation

Usage: python3 spuma-multiphysics_amd/profile_drivers.py stage 160 3
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from reactive_amd import abi, cases, lib as rlib, model  # noqa: E402


def stage(n, reps):
    lib = rlib.load()
    faces, volume, q, seeds, _ = cases.pure_n2o_cube(n, seed=1)
    st = np.zeros(len(q), dtype=abi.TSTATE_DTYPE)
    fluid = cases.PureFluid("N2O")
    st["p"] = np.maximum(seeds["p"], 1e3); st["T"] = seeds["T"]; st["rho"] = q[:, 0]; st["cv"] = 800.0
    st["sound"] = 200.0; st["gasMass"] = q[:, 0]
    cfg = abi.Config(cells=len(q), species=1, variables=5, faces=len(faces), fixed=0, viscosity=0, conductivity=0,
                     diffusivity=0, waveFactor=1.1, maxBytes=100e9, mechanical=0)
    err = C.create_string_buffer(1024)
    h = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(volume), rlib.ptr(faces), None, None, None, None,
                                      err, 1024)
    assert h, err.value
    b = np.zeros(5)
    for _ in range(reps):
        for s in (0, 1):
            rc = lib.reactive_transport_stage(h, rlib.ptr(q), rlib.ptr(st), None, None, 1e-9, s, rlib.ptr(b))
            assert rc == 0, lib.reactive_transport_error(h)
    lib.reactive_transport_destroy(h)
    print(f"stage: {len(q)} cells, {reps} SSPRK2 stage pairs", flush=True)
    del fluid


def flash(n):
    lib = rlib.load()
    faces, volume, q, seeds, _ = cases.pure_n2o_cube(n, seed=1)
    m = model.model_bytes(model.build_model())
    err = C.create_string_buffer(1024)
    h = lib.reactive_gpu_hem_create_v1(m, len(m), len(q), err, 1024)
    assert h, err.value
    e = cases.internal_energy(q, 1)
    s = seeds.copy()
    ok = np.zeros(len(q), np.int32)
    prof = abi.HemProfile.fresh()
    rc = lib.reactive_gpu_hem_run_v1(h, rlib.ptr(np.ascontiguousarray(q[:, :1])), rlib.ptr(e), len(q), rlib.ptr(s),
                                     rlib.ptr(ok), C.byref(prof), err, 1024)
    assert rc == 0, err.value
    print(f"flash: {len(q)} cells, {int(ok.sum())} converged, {prof.phaseEvaluations} phase evaluations, "
          f"kernel {prof.kernelSeconds:.4f} s", flush=True)
    lib.reactive_gpu_hem_destroy_v1(h)


def uniform(n, T, rho, mode, steps):
    lib = rlib.load()
    faces, volume = cases.cartesian_faces(n, 0.08 / n)
    nc = n ** 3
    e, pr, liquid, dome = cases.equilibrium_states(np.array([T]), np.array([rho]))
    q = np.zeros((nc, 5))
    q[:, 0] = rho
    q[:, 4] = rho * e[0]
    mb = model.model_bytes(model.build_model())
    err = C.create_string_buffer(1024)
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), nc, err, 1024)
    assert hem, err.value
    seed = np.zeros(1, dtype=abi.STATE_DTYPE)
    seed["p"] = pr; seed["T"] = T; seed["rho"] = rho; seed["liquidMass"][:, 0] = liquid
    seed["activeLiquids"] = int(liquid[0] > 0)
    ok = np.zeros(1, np.int32)
    prof = abi.HemProfile.fresh()
    qe = np.array([rho * e[0]])
    assert lib.reactive_gpu_hem_run_v1(hem, rlib.ptr(np.array([[rho]])), rlib.ptr(qe), 1, rlib.ptr(seed),
                                       rlib.ptr(ok), C.byref(prof), err, 1024) == 0 and ok[0] == 1, err.value
    states = np.repeat(seed, nc)
    cfg = abi.Config(cells=nc, species=1, variables=5, faces=len(faces), fixed=0, viscosity=0, conductivity=0,
                     diffusivity=0, waveFactor=1.1, maxBytes=100e9, mechanical=0)
    tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(volume), rlib.ptr(faces), None, None, None, None,
                                       err, 1024)
    step = lib.reactive_amd_step_create(tr, hem, err, 1024)
    assert step, err.value
    assert lib.reactive_amd_step_load(step, rlib.ptr(q), rlib.ptr(states)) == 0
    rep = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, 0.4, 1.0, steps, mode, C.byref(rep)) == 0, \
        lib.reactive_amd_step_error(step)
    print(f"uniform: {nc} cells at T={T} rho={rho}, mode {mode}, {steps} steps: FP64 phase evaluations "
          f"{rep.phaseEvaluations}, FP32 phase evaluations {rep.singlePhaseEvaluations}, flash failures "
          f"{rep.flashFailures[0]}/{rep.flashFailures[1]}, flash {rep.flashSeconds[0] + rep.flashSeconds[1]:.4f} s, "
          f"search {rep.searchSeconds:.4f} s", flush=True)
    lib.reactive_amd_step_destroy(step)
    lib.reactive_transport_destroy(tr)
    lib.reactive_gpu_hem_destroy_v1(hem)


def cube(n, mode, steps):
    """Resident SSPRK2 steps on the seeded pure-N2O N^3 box of the benchmark."""
    lib = rlib.load()
    faces, volume, q, seeds, _ = cases.pure_n2o_cube(n, seed=2026)
    nc = len(q)
    mb = model.model_bytes(model.build_model())
    err = C.create_string_buffer(1024)
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), nc, err, 1024)
    assert hem, err.value
    e = cases.internal_energy(q, 1)
    states = seeds.copy()
    ok = np.zeros(nc, np.int32)
    prof = abi.HemProfile.fresh()
    assert lib.reactive_gpu_hem_run_v1(hem, rlib.ptr(np.ascontiguousarray(q[:, :1])), rlib.ptr(e), nc,
                                       rlib.ptr(states), rlib.ptr(ok), C.byref(prof), err, 1024) == 0, err.value
    bad = np.where(ok == 0)[0]
    if bad.size:  # failed synthetic cells copy a converged neighbour
        good = np.where(ok == 1)[0]
        src = good[np.searchsorted(good, bad).clip(0, good.size - 1)]
        q[bad] = q[src]
        states[bad] = states[src]
    cfg = abi.Config(cells=nc, species=1, variables=5, faces=len(faces), fixed=0, viscosity=0, conductivity=0,
                     diffusivity=0, waveFactor=1.1, maxBytes=100e9, mechanical=0)
    tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(volume), rlib.ptr(faces), None, None, None, None,
                                       err, 1024)
    step = lib.reactive_amd_step_create(tr, hem, err, 1024)
    assert step, err.value
    assert lib.reactive_amd_step_load(step, rlib.ptr(q), rlib.ptr(states)) == 0
    rep = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, 0.4, 1.0, steps, mode, C.byref(rep)) == 0, \
        lib.reactive_amd_step_error(step)
    print(f"cube: {nc} cells, mode {mode}, {steps} steps: FP64 phase evaluations {rep.phaseEvaluations}, "
          f"FP32 phase evaluations {rep.singlePhaseEvaluations}, flash {rep.flashSeconds[0] + rep.flashSeconds[1]:.4f} s, "
          f"search {rep.searchSeconds:.4f} s", flush=True)
    lib.reactive_amd_step_destroy(step)
    lib.reactive_transport_destroy(tr)
    lib.reactive_gpu_hem_destroy_v1(hem)


def bench_flow(n, mode, steps):
    sys.path.insert(0, os.path.dirname(HERE))
    import bench
    lib = rlib.load()
    faces, volume, q, states, m, hem, info = bench.prepare(n, 2026)
    cfg = abi.Config(cells=len(q), species=1, variables=5, faces=len(faces), fixed=0, viscosity=0, conductivity=0,
                     diffusivity=0, waveFactor=1.1, maxBytes=200e9, mechanical=0)
    err = C.create_string_buffer(1024)
    tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(volume), rlib.ptr(faces), None, None, None, None,
                                       err, 1024)
    step = lib.reactive_amd_step_create(tr, hem, err, 1024)
    assert step, err.value
    rep, wall, kernels, _, _ = bench.run_steps(lib, step, q, states, 0.4, steps, mode, 2, 1)
    print(f"bench flow: {len(q)} cells, mode {mode}, {steps} timed steps, {wall / steps * 1e3:.2f} ms per step; "
          f"shapes: flash block {kernels.flashBlock}, state bins {kernels.stateBins}, tile depth {kernels.tileDepth}",
          flush=True)
    lib.reactive_amd_step_destroy(step)
    lib.reactive_transport_destroy(tr)
    lib.reactive_gpu_hem_destroy_v1(hem)


def copy(mib):
    import torch
    n = int(mib) << 20
    src = torch.ones(n // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)  # one elementwise copy kernel: reads and writes n bytes
    torch.cuda.synchronize()
    print(f"copy: 3 x {mib} MiB device-to-device (torch elementwise copy kernel)", flush=True)


if __name__ == "__main__":
    what = sys.argv[1]
    if what == "stage":
        stage(int(sys.argv[2]), int(sys.argv[3]))
    elif what == "flash":
        flash(int(sys.argv[2]))
    elif what == "copy":
        copy(int(sys.argv[2]))
    elif what == "bench":
        bench_flow(int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4]))
    elif what == "cube":
        cube(int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4]))
    elif what == "uniform":
        uniform(int(sys.argv[2]), float(sys.argv[3]), float(sys.argv[4]), int(sys.argv[5]), int(sys.argv[6]))
    else:
        raise SystemExit(__doc__)
