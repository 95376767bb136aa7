"""HEM UV flash on the GPU (reactive_gpu_hem_run_v1) against the oracle.

Two gates.  The device flash uses the elementary functions of
csrc/elementary.h; the oracle's twin build uses the same ones, so the device
result must equal it bit for bit (every field of every state, the success
flags and the phase evaluation count).  And the device flash must meet the
reference itself (the glibc oracle, equal to the compiled reference bit for
bit) at 1e-12: p, T, rho, e, c_f and the vapour fractions on every cell, c_eq
on every cell except those where a last-bit difference of glibc's elementary
functions moved the finite-difference Jacobian (helpers.assert_reference_parity)."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import GOLDEN, assert_reference_parity, cube_case, oracle_flash
from reactive_amd import abi, lib as rlib

pytestmark = pytest.mark.gpu

FIELDS = ("p", "T", "rho", "e", "entropy", "soundFrozen", "soundEquilibrium", "alphaGas", "gasMass", "cp", "cv")


def run_gpu(mb, q, e, seeds):
    lib = rlib.load()
    err = C.create_string_buffer(1024)
    h = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(e), err, 1024)
    assert h, err.value
    states = seeds.copy()
    success = np.zeros(len(e), np.int32)
    prof = abi.HemProfile.fresh()
    rc = lib.reactive_gpu_hem_run_v1(h, rlib.ptr(np.ascontiguousarray(q)), rlib.ptr(e), len(e), rlib.ptr(states),
                                     rlib.ptr(success), C.byref(prof), err, 1024)
    assert rc == 0, err.value
    lib.reactive_gpu_hem_destroy_v1(h)
    return states, success, prof


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def check_bitwise(mb, q, e, seeds, analytic=False):
    gs, gsucc, prof = run_gpu(mb, q, e, seeds)
    ts, tsucc, tcnt = oracle_flash(mb, q, e, seeds, twin=True)
    assert np.array_equal(gsucc, tsucc)
    rows = gs.view(np.uint8).reshape(len(e), -1) == ts.view(np.uint8).reshape(len(e), -1)
    bad = np.where(~rows.all(1))[0]
    assert bad.size == 0, f"{bad.size} cells differ from the twin oracle, first {bad[:8]}"
    assert prof.phaseEvaluations == int(tcnt[:, 0].sum())
    assert prof.residualEvaluations == int(tcnt[:, 1].sum())
    assert prof.deviceFailures == int((tsucc == 0).sum())
    assert prof.analyticJacobians == int(tcnt[:, 5].sum())
    assert prof.finiteDifferenceJacobians == int(tcnt[:, 6].sum())
    gl, glsucc, _ = oracle_flash(mb, q, e, seeds)
    d = assert_reference_parity(gs, gsucc, gl, glsucc, analytic=analytic)
    print(f"GPU vs the reference (glibc oracle), {d['cells']} cells, {d['not_bitwise'].size} not bit for bit:",
          {f: d[f][0] for f in ("p", "T", "soundFrozen", "soundEquilibrium", "alphaGas", "vapourFraction")})


def test_flash_cube_bitwise():
    case = cube_case(8, seed=1)
    check_bitwise(case["model"], case["q"][:, :1], case["energy"], case["seeds"])


def test_flash_edge_classes_bitwise():
    """Near-critical, compressed liquid, dilute vapour, the 148 K liquid limit
    and the liquid edge of the dome (the golden edge set)."""
    g = np.load(os.path.join(GOLDEN, "flash_n2o.npz"))
    check_bitwise(g["model"].tobytes(), g["edge_q"], g["edge_energy"], g["edge_seeds"])


def test_flash_binary_bitwise():
    """Binary N2O/N2 gas with liquid N2O: the two-species flash (golden inputs)."""
    g = np.load(os.path.join(GOLDEN, "flash_binary.npz"))
    check_bitwise(g["model"].tobytes(), g["q"], g["energy"], g["seeds"])


def test_flash_larger_cube_bitwise():
    case = cube_case(16, seed=5)
    check_bitwise(case["model"], case["q"][:, :1], case["energy"], case["seeds"])


CAP_DTYPE = np.dtype([("color", "<f8"), ("pressureJump", "<f8"), ("equilibrium", "<i4"), ("pad", "<i4")])


def test_resident_frozen_batch_matches_host_batch_and_oracle():
    """reactive_gpu_hem_run_resident_v1 with device pointers and a cell order,
    against reactive_gpu_hem_run_v2 (host arrays) and the oracle's
    frozen-inventory recovery (reactiveDeviceFlash.h:572-590), bit for bit."""
    from helpers import Device, oracle, p
    lib = rlib.load()
    case = cube_case(6, seed=17)
    mb, q, e, seeds = case["model"], np.ascontiguousarray(case["q"][:, :1]), case["energy"], case["states"]
    n = len(e)
    cap = np.zeros(n, CAP_DTYPE)
    # host batch (run_v2, equilibrium = 0)
    err = C.create_string_buffer(1024)
    h = lib.reactive_gpu_hem_create_v1(mb, len(mb), n, err, 1024)
    assert h, err.value
    hs, hsucc = seeds.copy(), np.zeros(n, np.int32)
    prof = abi.HemProfile.fresh()
    assert lib.reactive_gpu_hem_run_v2(h, rlib.ptr(q), rlib.ptr(e), rlib.ptr(cap), n, rlib.ptr(hs), rlib.ptr(hsucc),
                                       C.byref(prof), err, 1024) == 0, err.value
    lib.reactive_gpu_hem_destroy_v1(h)
    # resident batch over every other cell, in reverse order
    order = np.arange(n - 1, -1, -2, dtype=np.uint32)
    dev = Device()
    dq, de, ds, dc, do = dev.put(q), dev.put(e), dev.put(seeds), dev.put(cap), dev.put(order)
    blank = np.zeros(n, abi.HEM_OUTPUT_DTYPE)
    dout = dev.put(blank)
    r = lib.reactive_gpu_hem_create_resident_v1(mb, len(mb), n, err, 1024)
    assert r, err.value
    rprof = abi.HemProfile.fresh()
    rc = lib.reactive_gpu_hem_run_resident_v1(r, dq, de, ds, dc, do, len(order), dout, C.byref(rprof), err, 1024)
    assert rc == 0, err.value
    out = dev.get(dout, blank)
    assert rprof.submitted == len(order) and rprof.succeeded == int(out["success"][order].sum())
    untouched = np.setdiff1d(np.arange(n), order)
    assert not out.view(np.uint8).reshape(n, -1)[untouched].any()
    assert np.array_equal(out["success"][order], hsucc[order])
    good = order[hsucc[order] == 1]
    assert out["state"][good].tobytes() == hs[good].tobytes()
    # oracle frozen recovery, every submitted cell: the twin bit for bit, the
    # glibc build (the reference) at 1e-12
    img = C.create_string_buffer(mb, len(mb))
    ref = {}
    for twin in (True, False):
        L = oracle(twin)
        o = np.zeros(n, abi.STATE_DTYPE)
        okc = np.zeros(n, np.int32)
        for c in order:
            okc[c] = L.oracle_flash_frozen(img, p(q[c]), C.c_double(e[c]), p(seeds[c:c + 1]), p(o[c:c + 1]), None)
        ref[twin] = (o[order], okc[order])
    assert np.array_equal(out["success"][order], ref[True][1])
    good = out["success"][order] == 1
    assert out["state"][order][good].tobytes() == ref[True][0][good].tobytes()
    assert_reference_parity(out["state"][order], out["success"][order], ref[False][0], ref[False][1])
    # cells asking for the capillary equilibrium recovery and one with an
    # interface the reference rejects (a jump without a liquid is fine here;
    # color 2 is not): the first as the oracle recovers them, the last failed
    from helpers import oracle_capillary
    cap2 = cap.copy()
    cap2["equilibrium"][order[3]] = cap2["equilibrium"][order[5]] = 1
    cap2["color"][order[5]], cap2["pressureJump"][order[5]] = 0.4, 3e3
    cap2["color"][order[7]] = 2.0
    dc2 = dev.put(cap2)
    dout2 = dev.put(blank)
    rc = lib.reactive_gpu_hem_run_resident_v1(r, dq, de, ds, dc2, do, len(order), dout2, C.byref(rprof), err, 1024)
    assert rc == 0, err.value
    out2 = dev.get(dout2, blank)
    ts, tok, _ = oracle_capillary(mb, q[order], e[order], seeds[order], cap2["color"][order],
                                  cap2["pressureJump"][order], cap2["equilibrium"][order], twin=True)
    assert np.array_equal(out2["success"][order], tok) and tok[7] == 0
    good = tok == 1
    assert out2["state"][order][good].tobytes() == ts[good].tobytes()
    dev.close()
    lib.reactive_gpu_hem_destroy_v1(r)


def test_error_contracts():
    """The reference's argument checks (reactiveGpuHem.cuh:190-233): create
    rejects a wrong model size and a zero capacity with a NULL handle and a
    message; run rejects an empty batch, one above the capacity and a profile
    of another ABI with rc 1, and writes nothing; a cell no flash recovers is
    a per-cell failure (rc 0, success 0, its state untouched, counted in
    deviceFailures)."""
    lib = rlib.load()
    case = cube_case(3, seed=11)
    mb, q, e, seeds = case["model"], case["q"][:, :1].copy(), case["energy"].copy(), case["seeds"]
    err = C.create_string_buffer(1024)
    assert not lib.reactive_gpu_hem_create_v1(mb, len(mb) - 8, len(e), err, 1024)
    assert b"Invalid" in err.value
    assert not lib.reactive_gpu_hem_create_v1(mb, len(mb), 0, err, 1024)
    h = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(e), err, 1024)
    assert h, err.value
    states = seeds.copy()
    success = np.full(len(e), -7, np.int32)
    prof = abi.HemProfile.fresh()
    for count in (0, len(e) + 1):
        rc = lib.reactive_gpu_hem_run_v1(h, rlib.ptr(q), rlib.ptr(e), count, rlib.ptr(states), rlib.ptr(success),
                                         C.byref(prof), err, 1024)
        assert rc == 1 and b"capacity" in err.value
    bad = abi.HemProfile.fresh()
    bad.abiVersion = 2
    assert lib.reactive_gpu_hem_run_v1(h, rlib.ptr(q), rlib.ptr(e), len(e), rlib.ptr(states), rlib.ptr(success),
                                       C.byref(bad), err, 1024) == 1
    assert np.array_equal(states.view(np.uint8), seeds.view(np.uint8)) and (success == -7).all()
    e2 = e.copy()
    e2[5] = np.nan
    prof = abi.HemProfile.fresh()
    assert lib.reactive_gpu_hem_run_v1(h, rlib.ptr(q), rlib.ptr(e2), len(e), rlib.ptr(states), rlib.ptr(success),
                                       C.byref(prof), err, 1024) == 0, err.value
    assert success[5] == 0 and prof.deviceFailures >= 1
    assert np.array_equal(states[5:6].view(np.uint8), seeds[5:6].view(np.uint8))
    lib.reactive_gpu_hem_destroy_v1(h)


@pytest.mark.parametrize("part", ["cube", "edge", "binary"])
def test_flash_stable_gas_prune_bitwise(part):
    """The opt-in stable-gas pruning (stableGasPrune = 1, reactiveDeviceFlash.h:509-523)
    on the golden inputs: the device equals the twin oracle bit for bit, and the
    twin's glibc build equals the reference (test_oracle_golden.py)."""
    g = np.load(os.path.join(GOLDEN, "flash_prune.npz"))
    base = np.load(os.path.join(GOLDEN, "flash_binary.npz" if part == "binary" else "flash_n2o.npz"))
    key = "" if part == "binary" else f"{part}_"
    mb = (g["model_binary"] if part == "binary" else g["model"]).tobytes()
    check_bitwise(mb, base[f"{key}q"], base[f"{key}energy"], base[f"{key}seeds"])


@pytest.mark.parametrize("part", ["cube", "edge", "binary", "cube12"])
def test_flash_analytic_jacobian(part):
    """The reference's analytic-Jacobian mode (analyticJacobian = 1,
    reactiveDeviceFlashJacobian.h, its fastest published configuration): the
    device equals the twin oracle bit for bit (analytic-Jacobian count
    included, no finite-difference Jacobian), and every field of every cell,
    the equilibrium sound speed included, is within 1e-12 of the reference."""
    g = np.load(os.path.join(GOLDEN, "flash_analytic.npz"))
    if part == "cube12":
        q, e, seeds = g["cube12_q"], g["cube12_energy"], g["cube12_seeds"]
    else:
        base = np.load(os.path.join(GOLDEN, "flash_binary.npz" if part == "binary" else "flash_n2o.npz"))
        key = "" if part == "binary" else f"{part}_"
        q, e, seeds = base[f"{key}q"], base[f"{key}energy"], base[f"{key}seeds"]
    mb = (g["model_binary"] if part == "binary" else g["model"]).tobytes()
    check_bitwise(mb, q, e, seeds, analytic=True)
    # and the reference's own outputs (the fixture): within the same gate
    gs, gsucc, prof = run_gpu(mb, q, e, seeds)
    assert prof.analyticJacobians > 0 and prof.finiteDifferenceJacobians == 0
    assert_reference_parity(gs, gsucc, g[f"{part}_out"], g[f"{part}_success"], analytic=True)
