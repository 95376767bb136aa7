"""The driver's attempt loop around a resident step (REACTIVE_AMD_STEP_RETRY,
ReactiveFoam.C:1443-1494): an attempt with a failed flash or a failed
post-stage-0 CFL recheck is rolled back and repeated with half its dt.

* A pure-N2O Riemann box at CFL 0.5 fails the recheck once (the oracle
  transport and flash show it: the stable step after stage 0 is below dt and
  equals dt / 2 after the halved attempt); the retried step must equal, bit
  for bit, a plain step whose maximum is the halved dt.
* A cell whose energy no flash can recover fails every attempt: after 12
  retries the call fails and the step-start q and states are back in place.
* In a batch only the failing case halves its dt; the other case's step is
  that of the case alone."""
import ctypes as C

import numpy as np
import pytest

from reactive_amd import abi, cases, lib as rlib, model
from test_gpu_batch import prepared, transport

pytestmark = pytest.mark.gpu

RETRY = 0x400  # REACTIVE_AMD_STEP_RETRY


def state_bytes(s):
    return s.view(np.uint8).reshape(len(s), -1)


def run(lib, hem, trs, q, s, mode, cfl=0.5, maximum=1.0):
    err = C.create_string_buffer(1024)
    if len(trs) == 1:
        step = lib.reactive_amd_step_create(trs[0], hem, err, 1024)
    else:
        step = lib.reactive_amd_step_create_batch((C.c_void_p * len(trs))(*trs), len(trs), hem, err, 1024)
    assert step, err.value
    assert lib.reactive_amd_step_load(step, rlib.ptr(q), rlib.ptr(s)) == 0
    rep = abi.StepReport.fresh()
    rc = lib.reactive_amd_step_run(step, cfl, maximum, 1, mode, C.byref(rep))
    message = lib.reactive_amd_step_error(step)
    qo, so = np.zeros_like(q), np.zeros_like(s)
    assert lib.reactive_amd_step_read(step, rlib.ptr(qo), rlib.ptr(so)) == 0
    boxes = []
    for i in range(len(trs)):
        r = abi.StepReport.fresh()
        if rc == 0:
            assert lib.reactive_amd_step_report_box(step, i, C.byref(r)) == 0
        boxes.append(r)
    lib.reactive_amd_step_destroy(step)
    return rc, message, rep, qo, so, boxes


def setup(lib, boxes):
    mb = model.model_bytes(model.build_model())
    err = C.create_string_buffer(1024)
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), sum(len(b[2]) for b in boxes), err, 1024)
    assert hem, err.value
    prepared_boxes = [prepared(lib, hem, b + (None,)) for b in boxes]
    trs = [transport(lib, b, 0) for b in boxes]
    return hem, trs, prepared_boxes


@pytest.mark.parametrize("mode", [0, 1, 0x800], ids=["fp64", "mixed", "fp64-reuse"])
def test_retried_step_equals_step_at_half_dt(mode):
    lib = rlib.load()
    box = cases.riemann_cube(12)
    hem, trs, [(q, s)] = setup(lib, [box])
    plain = run(lib, hem, trs, q, s, mode)
    assert plain[0] == 0 and plain[2].cflViolations == 1 and plain[2].retries == 0
    rc, message, rep, qr, sr, _ = run(lib, hem, trs, q, s, mode | RETRY)
    print(f"\nplain dt {plain[2].dt:.6e}, retried dt {rep.dt:.6e}, retries {rep.retries}", flush=True)
    assert rc == 0, message
    assert rep.retries == 1
    assert rep.dt == 0.5 * plain[2].dt
    rc, message, rh, qh, sh, _ = run(lib, hem, trs, q, s, mode, maximum=rep.dt)
    assert rc == 0, message
    assert rh.dt == rep.dt and rh.cflViolations == 0
    assert np.array_equal(qr.view(np.uint64), qh.view(np.uint64))
    assert np.array_equal(state_bytes(sr), state_bytes(sh))
    assert np.array_equal(np.array(rep.totals[:5]), np.array(rh.totals[:5]))
    for t in trs:
        lib.reactive_transport_destroy(t)
    lib.reactive_gpu_hem_destroy_v1(hem)


def test_unrecoverable_cell_restores_the_step_start():
    lib = rlib.load()
    box = cases.riemann_cube(8)
    hem, trs, [(q, s)] = setup(lib, [box])
    q = q.copy()
    q[100, 4] *= 1e3  # an energy far above the model's temperature range: no flash recovers it
    rc, _, rep, _, _, _ = run(lib, hem, trs, q, s, 0)
    assert rc == 0 and rep.flashFailures[0] >= 1  # without retries: counted, the cell keeps its state
    rc, message, _, qo, so, _ = run(lib, hem, trs, q, s, RETRY)
    print("\n" + message.decode(), flush=True)
    assert rc == 1 and b"13 attempts" in message
    assert np.array_equal(qo.view(np.uint64), q.view(np.uint64))
    assert np.array_equal(state_bytes(so), state_bytes(s))
    for t in trs:
        lib.reactive_transport_destroy(t)
    lib.reactive_gpu_hem_destroy_v1(hem)


def test_batch_retries_only_the_failing_case():
    lib = rlib.load()
    quiet = cases.riemann_cube(10, rho_left=20.0, rho_right=20.0)  # uniform: no wave, no retry
    boxes = [cases.riemann_cube(12), quiet]
    hem, trs, prepared_boxes = setup(lib, boxes)
    qs, ss = zip(*prepared_boxes)
    q0, s0 = np.concatenate(qs), np.concatenate(ss)
    rc, message, rep, qb, sb, per_box = run(lib, hem, trs, q0, s0, RETRY)
    assert rc == 0, message
    assert rep.retries == 1
    offset = 0
    for i, (q, s) in enumerate(prepared_boxes):
        rc, message, ra, qa, sa, _ = run(lib, hem, [trs[i]], q, s, RETRY)
        assert rc == 0, message
        assert ra.retries == (1 if i == 0 else 0)
        assert per_box[i].dt == ra.dt
        n = len(q)
        assert np.array_equal(qb[offset:offset + n].view(np.uint64), qa.view(np.uint64))
        assert np.array_equal(state_bytes(sb[offset:offset + n]), state_bytes(sa))
        offset += n
    for t in trs:
        lib.reactive_transport_destroy(t)
    lib.reactive_gpu_hem_destroy_v1(hem)
