"""Exact reuse within a flash (REACTIVE_AMD_STEP_REUSE; the reference pool's
exactBatchReuse, reactiveClosurePool.inc:3-54): cells with bitwise-equal
flash inputs share the flash of the lowest such cell.  On a Riemann box (two
uniform halves) the steps must equal those that flash every cell, bit for bit,
in both flash arithmetics and from captured graphs, while most cells reuse."""
import ctypes as C

import numpy as np
import pytest

from reactive_amd import abi, cases, lib as rlib, model
from test_gpu_batch import prepared, transport

pytestmark = pytest.mark.gpu

REUSE = 0x800  # REACTIVE_AMD_STEP_REUSE
GRAPH = 0x100  # REACTIVE_AMD_STEP_GRAPH


def steps(lib, hem, tr, q, s, mode, count):
    err = C.create_string_buffer(1024)
    step = lib.reactive_amd_step_create(tr, hem, err, 1024)
    assert step, err.value
    assert lib.reactive_amd_step_load(step, rlib.ptr(q), rlib.ptr(s)) == 0
    rep = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, 0.4, 1.0, count, mode, C.byref(rep)) == 0, lib.reactive_amd_step_error(step)
    qo, so = np.zeros_like(q), np.zeros_like(s)
    assert lib.reactive_amd_step_read(step, rlib.ptr(qo), rlib.ptr(so)) == 0
    lib.reactive_amd_step_destroy(step)
    return rep, qo, so


@pytest.mark.parametrize("mode", [0, 1, GRAPH], ids=["fp64", "mixed", "fp64-graph"])
def test_reuse_equals_flashing_every_cell(mode):
    lib = rlib.load()
    box = cases.riemann_cube(16)
    mb = model.model_bytes(model.build_model())
    err = C.create_string_buffer(1024)
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(box[2]), err, 1024)
    assert hem, err.value
    q, s = prepared(lib, hem, box + (None,))
    tr = transport(lib, box, 0)
    count = 3
    plain, qp, sp = steps(lib, hem, tr, q, s, mode, count)
    reused, qr, sr = steps(lib, hem, tr, q, s, mode | REUSE, count)
    flashes = 2 * count
    print(f"\ncells {len(q)}, reused {reused.reusedCells} of {flashes * len(q)} cell flashes; FP64 phase "
          f"evaluations {plain.phaseEvaluations} -> {reused.phaseEvaluations}", flush=True)
    assert np.array_equal(qr.view(np.uint64), qp.view(np.uint64))
    assert np.array_equal(sr.view(np.uint8).reshape(len(sr), -1), sp.view(np.uint8).reshape(len(sp), -1))
    assert np.array_equal(np.array(reused.totals[:5]), np.array(plain.totals[:5]))
    assert plain.reusedCells == 0
    assert reused.reusedCells > 0.8 * flashes * len(q)
    assert reused.phaseEvaluations < 0.3 * plain.phaseEvaluations
    lib.reactive_transport_destroy(tr)
    lib.reactive_gpu_hem_destroy_v1(hem)
