"""Batch of independent cases (reactive_amd_step_create_batch): uneven
pure-N2O cubes stepped together, one flux launch per SSPRK2 stage and one
flash over every cell, each with its own CFL dt.  Every case must equal the
same case stepped alone on the device, bit for bit (q, every state field,
dt, recheck step, totals, boundary integral)."""
import ctypes as C

import numpy as np
import pytest

from helpers import config
from reactive_amd import abi, cases, lib as rlib, model

pytestmark = pytest.mark.gpu


def prepared(lib, hem, case):
    """Flash-consistent step-start states from the device flash; failed cells
    take the conserved values and state of a converged cell."""
    faces, volume, q, seeds, _ = case
    err = C.create_string_buffer(1024)
    states = seeds.copy()
    ok = np.zeros(len(q), np.int32)
    prof = abi.HemProfile.fresh()
    assert lib.reactive_gpu_hem_run_v1(hem, rlib.ptr(np.ascontiguousarray(q[:, :1])),
                                       rlib.ptr(cases.internal_energy(q, 1)), len(q), rlib.ptr(states), rlib.ptr(ok),
                                       C.byref(prof), err, 1024) == 0, err.value
    good, bad = np.where(ok == 1)[0], np.where(ok == 0)[0]
    if bad.size:
        q[bad] = q[good[0]]
        states[bad] = states[good[0]]
    return q, states


def transport(lib, case, scheme):
    faces, volume = case[0], case[1]
    err = C.create_string_buffer(1024)
    tr = lib.reactive_transport_create(1, C.byref(config(len(volume), len(faces))), rlib.ptr(volume),
                                       rlib.ptr(faces), None, None, None, None, err, 1024)
    assert tr, err.value
    assert lib.reactive_amd_transport_set_flux(tr, scheme) == 0
    return tr


@pytest.mark.parametrize("mode,scheme", [(0, 0), (1, 0), (0, 1), (0 | 0x100, 0)],
                         ids=["fp64-hll", "mixed-hll", "fp64-hllc", "fp64-graph"])
def test_batch_equals_cases_alone(mode, scheme):
    lib = rlib.load()
    batch = cases.multi_case(6, seed=7, edges=(4, 6, 8, 10, 12))
    sizes = [len(c[2]) for c in batch]
    print(f"batch edges {[round(s ** (1 / 3)) for s in sizes]}, mode {mode:#x}, scheme {scheme}", flush=True)
    mb = model.model_bytes(model.build_model())
    err = C.create_string_buffer(1024)
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), sum(sizes), err, 1024)
    assert hem, err.value
    qs, ss = zip(*[prepared(lib, hem, c) for c in batch])
    trs = [transport(lib, c, scheme) for c in batch]
    handles = (C.c_void_p * len(trs))(*trs)
    step = lib.reactive_amd_step_create_batch(handles, len(trs), hem, err, 1024)
    assert step, err.value
    q0, s0 = np.concatenate(qs), np.concatenate(ss)
    assert lib.reactive_amd_step_load(step, rlib.ptr(q0), rlib.ptr(s0)) == 0
    steps = 3 if mode & 0x100 else 2
    rep = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, 0.4, 1.0, steps, mode, C.byref(rep)) == 0, lib.reactive_amd_step_error(step)
    qb = np.zeros_like(q0)
    sb = np.zeros(len(q0), dtype=abi.STATE_DTYPE)
    assert lib.reactive_amd_step_read(step, rlib.ptr(qb), rlib.ptr(sb)) == 0
    boxes = []
    for i in range(len(trs)):
        r = abi.StepReport.fresh()
        assert lib.reactive_amd_step_report_box(step, i, C.byref(r)) == 0
        boxes.append(r)
    assert lib.reactive_amd_step_report_box(step, len(trs), C.byref(abi.StepReport.fresh())) == 1
    lib.reactive_amd_step_destroy(step)
    dts = []
    offset = 0
    failures = 0
    for i, tr in enumerate(trs):
        single = lib.reactive_amd_step_create(tr, hem, err, 1024)
        assert single, err.value
        assert lib.reactive_amd_step_load(single, rlib.ptr(qs[i]), rlib.ptr(ss[i])) == 0
        r = abi.StepReport.fresh()
        assert lib.reactive_amd_step_run(single, 0.4, 1.0, steps, mode, C.byref(r)) == 0, \
            lib.reactive_amd_step_error(single)
        q = np.zeros_like(qs[i])
        s = np.zeros(len(q), dtype=abi.STATE_DTYPE)
        assert lib.reactive_amd_step_read(single, rlib.ptr(q), rlib.ptr(s)) == 0
        lib.reactive_amd_step_destroy(single)
        n = len(q)
        assert np.array_equal(qb[offset:offset + n], q), f"case {i}: q differs"
        assert sb[offset:offset + n].tobytes() == s.tobytes(), f"case {i}: states differ"
        b = boxes[i]
        assert b.dt == r.dt and b.recheckDt == r.recheckDt and b.cflViolations == r.cflViolations
        assert b.steps == steps
        assert list(b.totals)[:5] == list(r.totals)[:5]
        assert list(b.boundaryIntegral)[:5] == list(r.boundaryIntegral)[:5]
        dts.append(r.dt)
        failures += r.flashFailures[0] + r.flashFailures[1]
        offset += n
    assert len(set(dts)) > 1  # the cases really step with their own dt
    assert rep.dt == min(dts)
    assert rep.flashFailures[0] + rep.flashFailures[1] == failures
    for t in trs:
        lib.reactive_transport_destroy(t)
    lib.reactive_gpu_hem_destroy_v1(hem)


def test_batch_rejects_mixed_inputs():
    """Creation contract: null or empty lists, and boxes with different
    species or flux schemes, are refused with a message."""
    lib = rlib.load()
    err = C.create_string_buffer(1024)
    mb = model.model_bytes(model.build_model())
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), 1000, err, 1024)
    assert not lib.reactive_amd_step_create_batch(None, 2, hem, err, 1024) and err.value
    batch = cases.multi_case(2, seed=3, edges=(4, 6))
    trs = [transport(lib, c, 0) for c in batch]
    assert lib.reactive_amd_transport_set_flux(trs[1], 1) == 0
    handles = (C.c_void_p * 2)(*trs)
    assert not lib.reactive_amd_step_create_batch(handles, 2, hem, err, 1024)
    assert b"flux" in err.value
    assert not lib.reactive_amd_step_create_batch(handles, 0, hem, err, 1024)
    for t in trs:
        lib.reactive_transport_destroy(t)
    lib.reactive_gpu_hem_destroy_v1(hem)
