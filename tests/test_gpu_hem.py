"""HEM UV flash on the GPU (reactive_gpu_hem_run_v1) against the oracle.

The device flash uses the elementary functions of csrc/elementary.h; the
oracle's twin build uses the same ones, so the device result must equal it
bit for bit (every field of every state, the success flags and the phase
evaluation count).  The distance to the glibc build (which equals the
compiled reference) is the twin-vs-glibc envelope, bounded on the CPU in
test_twin_envelope.py; it is printed here for the record."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import GOLDEN, cube_case, oracle_flash
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


def check_bitwise(mb, q, e, seeds):
    gs, gsucc, prof = run_gpu(mb, q, e, seeds)
    ts, tsucc, tcnt = oracle_flash(mb, q, e, seeds, twin=True)
    assert np.array_equal(gsucc, tsucc)
    rows = gs.view(np.uint8).reshape(len(e), -1) == ts.view(np.uint8).reshape(len(e), -1)
    bad = np.where(~rows.all(1))[0]
    assert bad.size == 0, f"{bad.size} cells differ from the twin oracle, first {bad[:8]}"
    assert prof.phaseEvaluations == int(tcnt[:, 0].sum())
    assert prof.residualEvaluations == int(tcnt[:, 1].sum())
    assert prof.deviceFailures == int((tsucc == 0).sum())
    gl, glsucc, _ = oracle_flash(mb, q, e, seeds)
    ok = (glsucc == 1) & (gsucc == 1)
    print("GPU vs glibc oracle, max relative difference:",
          {f: float(rel(gs[f][ok], gl[f][ok]).max()) for f in FIELDS})


def test_flash_cube_bitwise():
    case = cube_case(8, seed=1)
    check_bitwise(case["model"], case["q"][:, :1], case["energy"], case["seeds"])


def test_flash_edge_classes_bitwise():
    """Near-critical, compressed liquid, dilute vapour, the 148 K liquid limit
    and the liquid edge of the dome (the golden edge set)."""
    g = np.load(os.path.join(GOLDEN, "flash_n2o.npz"))
    check_bitwise(g["model"].tobytes(), g["edge_q"], g["edge_energy"], g["edge_seeds"])


def test_flash_larger_cube_bitwise():
    case = cube_case(16, seed=5)
    check_bitwise(case["model"], case["q"][:, :1], case["energy"], case["seeds"])
