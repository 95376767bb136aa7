"""Batch of independent cases (reactive_amd_step_create_batch): uneven
pure-N2O cubes stepped together, one flux launch per SSPRK2 stage and one
flash over every cell, each with its own CFL dt.  Every case must equal the
same case stepped alone on the device, bit for bit (q, every state field,
dt, recheck step, totals, boundary integral).  At full size (96 cubes with
edges drawn from {16, 24, 32, 40, 48} cells, about 4.3 million cells) three
of the cubes are also stepped on the oracle: bit for bit with its twin build
and at the parity gate with the reference's arithmetic."""
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


def test_batch_96_uneven_cubes_full_size():
    """The uneven multi-case workload at its size: 96 pure-N2O cubes, edges
    uniform over {16, 24, 32, 40, 48} (2 mm cells, each cube seeded as the
    40^3 box), one FP64 step of the whole batch.  Every cube equals its own
    single-case step bit for bit; three cubes (the first of each of three
    sizes) equal the oracle's step."""
    from test_gpu_step import check_reference_step, oracle_step
    lib = rlib.load()
    batch = cases.multi_case(96, seed=2026)
    sizes = [len(c[2]) for c in batch]
    edges = [round(n ** (1 / 3)) for n in sizes]
    print(f"96 cubes, {sum(sizes)} cells, edges {sorted(set(edges))}", flush=True)
    assert sum(sizes) > 4_000_000
    mb = model.model_bytes(model.build_model())
    err = C.create_string_buffer(1024)
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), sum(sizes), err, 1024)
    assert hem, err.value
    # flash-consistent step-start states, one device flash over every cube
    q_all = np.concatenate([c[2] for c in batch])
    seeds_all = np.concatenate([c[3] for c in batch])
    q_all, s_all = prepared(lib, hem, (None, None, q_all, seeds_all, None))
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    qs = [q_all[offsets[i]:offsets[i + 1]] for i in range(len(batch))]
    ss = [s_all[offsets[i]:offsets[i + 1]] for i in range(len(batch))]
    trs = [transport(lib, c, 0) for c in batch]
    handles = (C.c_void_p * len(trs))(*trs)
    step = lib.reactive_amd_step_create_batch(handles, len(trs), hem, err, 1024)
    assert step, err.value
    assert lib.reactive_amd_step_load(step, rlib.ptr(q_all), rlib.ptr(s_all)) == 0
    rep = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, 0.4, 1.0, 1, 0, C.byref(rep)) == 0, lib.reactive_amd_step_error(step)
    qb = np.zeros_like(q_all)
    sb = np.zeros(len(q_all), dtype=abi.STATE_DTYPE)
    assert lib.reactive_amd_step_read(step, rlib.ptr(qb), rlib.ptr(sb)) == 0
    boxes = []
    for i in range(len(trs)):
        r = abi.StepReport.fresh()
        assert lib.reactive_amd_step_report_box(step, i, C.byref(r)) == 0
        boxes.append(r)
    lib.reactive_amd_step_destroy(step)
    print(f"batch step {rep.seconds:.3f} s on the device, {rep.flashFailures[1]} failed second flashes", flush=True)
    singles = {}
    for i, tr in enumerate(trs):
        single = lib.reactive_amd_step_create(tr, hem, err, 1024)
        assert single, err.value
        assert lib.reactive_amd_step_load(single, rlib.ptr(qs[i]), rlib.ptr(ss[i])) == 0
        r = abi.StepReport.fresh()
        assert lib.reactive_amd_step_run(single, 0.4, 1.0, 1, 0, C.byref(r)) == 0, lib.reactive_amd_step_error(single)
        q = np.zeros_like(qs[i])
        s = np.zeros(len(q), dtype=abi.STATE_DTYPE)
        assert lib.reactive_amd_step_read(single, rlib.ptr(q), rlib.ptr(s)) == 0
        lib.reactive_amd_step_destroy(single)
        a, b = offsets[i], offsets[i + 1]
        assert np.array_equal(qb[a:b], q), f"cube {i}: q differs"
        assert sb[a:b].tobytes() == s.tobytes(), f"cube {i}: states differ"
        assert boxes[i].dt == r.dt and boxes[i].recheckDt == r.recheckDt
        assert list(boxes[i].totals)[:5] == list(r.totals)[:5]
        singles[i] = (q, s, r)
    assert len({b.dt for b in boxes}) > 1
    # three cubes of three sizes against the oracle's step
    picked = []
    for edge in sorted(set(edges))[:3]:
        picked.append(edges.index(edge))
    for i in picked:
        faces, volume = batch[i][0], batch[i][1]
        case = dict(model=mb, faces=faces, volume=volume, q=qs[i], states=ss[i], ns=1)
        ref = oracle_step(case, 0.4)
        q, s, r = singles[i]
        assert r.dt == ref["dt"]
        assert np.array_equal(q, ref["q"]), f"cube {i} ({edges[i]}^3): q differs from the oracle"
        assert s.tobytes() == ref["s2"].tobytes(), f"cube {i} ({edges[i]}^3): states differ from the oracle"
        check_reference_step(case, 0.4, 0, r, q, s)
        print(f"cube {i} ({edges[i]}^3) equals the oracle step", flush=True)
    for t in trs:
        lib.reactive_transport_destroy(t)
    lib.reactive_gpu_hem_destroy_v1(hem)
