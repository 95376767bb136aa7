"""The allocation-free resident HEM batch
(reactive_gpu_hem_run_resident_noalloc_v1): the same frozen-inventory
recovery as reactive_gpu_hem_run_resident_v1, its work counters in the
library's static device scratch.  Two calls in a row, the second with fewer
cells, each equal to run_resident_v1 on the same inputs bit for bit (outputs
and every profile counter); a batch beyond the scratch's 1,048,576 cells is
refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from helpers import Device, cube_case
from reactive_amd import abi, lib as rlib

pytestmark = pytest.mark.gpu

CAP_DTYPE = np.dtype([("color", "<f8"), ("pressureJump", "<f8"), ("equilibrium", "<i4"), ("pad", "<i4")])
COUNTERS = ("submitted", "succeeded", "deviceFailures", "phaseEvaluations", "residualEvaluations",
            "flashCandidates", "stableCandidates", "analyticJacobians", "finiteDifferenceJacobians")


def run(lib, entry, h, dev, bufs, order, n):
    dq, de, ds, dc = bufs
    do = dev.put(order)
    blank = np.zeros(n, abi.HEM_OUTPUT_DTYPE)
    dout = dev.put(blank)
    prof = abi.HemProfile.fresh()
    err = C.create_string_buffer(1024)
    rc = getattr(lib, entry)(h, dq, de, ds, dc, do, len(order), dout, C.byref(prof), err, 1024)
    assert rc == 0, err.value
    return dev.get(dout, blank), prof


def test_noalloc_resident_batch_equals_resident_batch():
    lib = rlib.load()
    # both entry points carry their ctypes signatures (a pointer passed
    # without them would be cut to 32 bits)
    for name in ("reactive_gpu_hem_run_resident_v1", "reactive_gpu_hem_run_resident_noalloc_v1"):
        assert getattr(lib, name).argtypes and len(getattr(lib, name).argtypes) == 11, name
    case = cube_case(12, seed=23)
    mb, q, e = case["model"], np.ascontiguousarray(case["q"][:, :1]), case["energy"]
    seeds = case["states"]
    n = len(e)
    cap = np.zeros(n, CAP_DTYPE)
    err = C.create_string_buffer(1024)
    h = lib.reactive_gpu_hem_create_resident_v1(mb, len(mb), n, err, 1024)
    assert h, err.value
    dev = Device()
    bufs = (dev.put(q), dev.put(e), dev.put(seeds), dev.put(cap))
    # two calls in a row: every cell in reverse order, then every third cell
    for order in (np.arange(n - 1, -1, -1, dtype=np.uint32), np.arange(0, n, 3, dtype=np.uint32)):
        print(f"batch of {len(order)} cells", flush=True)
        ref, rprof = run(lib, "reactive_gpu_hem_run_resident_v1", h, dev, bufs, order, n)
        out, prof = run(lib, "reactive_gpu_hem_run_resident_noalloc_v1", h, dev, bufs, order, n)
        assert out.tobytes() == ref.tobytes()
        for f in COUNTERS:
            assert getattr(prof, f) == getattr(rprof, f), f
        assert prof.succeeded > 0 and prof.phaseEvaluations > 0
    # a batch beyond the scratch is refused with a message, before any device work
    big = 1_048_577
    hb = lib.reactive_gpu_hem_create_resident_v1(mb, len(mb), big, err, 1024)
    assert hb, err.value
    prof = abi.HemProfile.fresh()
    rc = lib.reactive_gpu_hem_run_resident_noalloc_v1(hb, *bufs, dev.put(np.zeros(1, np.uint32)), big,
                                                      dev.put(np.zeros(1, abi.HEM_OUTPUT_DTYPE)), C.byref(prof), err,
                                                      1024)
    assert rc == 1 and b"1048576" in err.value
    lib.reactive_gpu_hem_destroy_v1(hb)
    dev.close()
    lib.reactive_gpu_hem_destroy_v1(h)
