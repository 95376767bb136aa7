#!/usr/bin/env python3
"""Writes HEM capture fixtures with the reference's own capture tool.

oracle/_ref/libhem_capture.so is tools/hem_capture.cpp compiled from the
reference's source (`make -C oracle capture`, this container only); it
forwards to oracle/_build/libcapture_target.so, a target without a flash,
and records the first batch of at least 1000 cells in the capture layout.
Each capture here is a seeded cube's equilibrium-flash batch (run_v1).  The
reference's outputs for the same batch (oracle/_ref, its compiled device
flash) are stored beside it.

    python tests/golden/make_capture.py          # both fixtures
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "spuma-multiphysics_amd"))
from reactive_amd import cases, model  # noqa: E402

CAPTURE_TOOL = os.path.join(ROOT, "oracle", "_ref", "libhem_capture.so")
TARGET = os.path.join(ROOT, "oracle", "_build", "libcapture_target.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref.so")

# name: (gas species, cube edge, seed)
SETS = {"capture_n2o": (("N2O",), 12, 61), "capture_n2o_n2": (("N2O", "N2"), 10, 62)}


def inputs(name):
    """The batch of a fixture: (model bytes, q[count, ns], energy, seeds)."""
    species, n, seed = SETS[name]
    if len(species) == 1:
        _, _, q, seeds, _ = cases.pure_n2o_cube(n, seed=seed)
    else:
        _, _, q, seeds, _ = cases.binary_cube(n, seed=seed)
    ns = len(species)
    mb = model.model_bytes(model.build_model(gas_species=species))
    return mb, np.ascontiguousarray(q[:, :ns]), cases.internal_energy(q, ns), seeds


def capture_one(name):
    """Runs in a process of its own: the tool records once per process."""
    mb, q, e, seeds = inputs(name)
    lib = C.CDLL(CAPTURE_TOOL)
    V = C.c_void_p
    lib.reactive_gpu_hem_create_v1.restype = V
    lib.reactive_gpu_hem_create_v1.argtypes = [V, C.c_size_t, C.c_size_t, C.c_char_p, C.c_size_t]
    lib.reactive_gpu_hem_run_v1.argtypes = [V, V, V, C.c_size_t, V, V, V, C.c_char_p, C.c_size_t]
    lib.reactive_gpu_hem_destroy_v1.argtypes = [V]
    err = C.create_string_buffer(1024)
    h = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(e), err, 1024)
    assert h, err.value
    states = seeds.copy()
    success = np.zeros(len(e), np.int32)
    profile = (C.c_uint8 * 144)()
    rc = lib.reactive_gpu_hem_run_v1(h, q.ctypes.data, e.ctypes.data, len(e), states.ctypes.data,
                                     success.ctypes.data, profile, err, 1024)
    assert rc == 0, err.value
    lib.reactive_gpu_hem_destroy_v1(h)


def reference_outputs(name):
    mb, q, e, seeds = inputs(name)
    lib = C.CDLL(REF)
    V = C.c_void_p
    lib.ref_flash_batch.argtypes = [V, C.c_size_t, V, V, V, C.c_int, V, V, V]
    out = seeds.copy()
    success = np.zeros(len(e), np.int32)
    counters = np.zeros((len(e), 9), np.uint64)
    lib.ref_flash_batch(C.create_string_buffer(mb, len(mb)), len(e), q.ctypes.data, e.ctypes.data,
                        seeds.ctypes.data, 0, out.ctypes.data, success.ctypes.data, counters.ctypes.data)
    np.savez_compressed(os.path.join(HERE, f"{name}_out.npz"), out=out, success=success, counters=counters)
    print(name, "reference success", int(success.sum()), "/", len(e))


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--one":
        capture_one(sys.argv[2])
        return
    for name, (species, _, _) in SETS.items():
        path = os.path.join(HERE, f"{name}.bin")
        if os.path.exists(path):
            os.remove(path)
        env = dict(os.environ, REACTIVE_CAPTURE_HEM_LIBRARY=TARGET, REACTIVE_CAPTURE_FILE=path,
                   REACTIVE_CAPTURE_SPECIES=str(len(species)))
        subprocess.run([sys.executable, os.path.abspath(__file__), "--one", name], env=env, check=True)
        assert os.path.exists(path), f"the capture tool wrote no {path}"
        reference_outputs(name)


if __name__ == "__main__":
    main()
