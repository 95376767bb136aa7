"""The capillary / boundary-search flash on the GPU (Flash<..., CAP> in
csrc/flash_device.h, hemFlashKernel<..., CAP>): capillary recoveries with a
fixed interface (reactive_gpu_hem_run_v2 from host arrays and
run_resident_v1 from device pointers) and boundaryRecovery = 1 models
(reactive_gpu_hem_run_v1), on the reference-written fixtures
(tests/golden/flash_capillary.npz).  Bit for bit equal to the oracle's twin
build (states, success flags, phase evaluations); at the FP64 parity gate
against the reference itself (the stored outputs of its compiled flash),
with up to 6 % of the cells not bit for bit with it (see check())."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import GOLDEN, Device, assert_reference_parity, oracle_capillary, oracle_flash
from reactive_amd import abi, lib as rlib
from reactive_amd.abi import MIXED

pytestmark = pytest.mark.gpu

CAPILLARY = ["n2o", "trace_capillary"]
BOUNDARY = ["trace", "trace_analytic", "binary"]


def load(name):
    z = np.load(os.path.join(GOLDEN, "flash_capillary.npz"))
    return {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(name + "__")}


def cap_records(d):
    cap = np.zeros(len(d["energy"]), abi.CAPILLARY_DTYPE)
    cap["color"], cap["pressureJump"], cap["equilibrium"] = d["color"], d["jump"], d["equilibrium"]
    return cap


def host_batch(d, capillary):
    lib = rlib.load()
    mb = d["model"].tobytes()
    n = len(d["energy"])
    err = C.create_string_buffer(1024)
    h = lib.reactive_gpu_hem_create_v1(mb, len(mb), n, err, 1024)
    assert h, err.value
    states, ok = d["seeds"].copy(), np.zeros(n, np.int32)
    prof = abi.HemProfile.fresh()
    q = np.ascontiguousarray(d["q"])
    if capillary:
        rc = lib.reactive_gpu_hem_run_v2(h, rlib.ptr(q), rlib.ptr(d["energy"]), rlib.ptr(cap_records(d)), n,
                                         rlib.ptr(states), rlib.ptr(ok), C.byref(prof), err, 1024)
    else:
        rc = lib.reactive_gpu_hem_run_v1(h, rlib.ptr(q), rlib.ptr(d["energy"]), n, rlib.ptr(states), rlib.ptr(ok),
                                         C.byref(prof), err, 1024)
    assert rc == 0, err.value
    lib.reactive_gpu_hem_destroy_v1(h)
    return states, ok, prof


def twin(d):
    mb = d["model"].tobytes()
    if "color" in d:
        return oracle_capillary(mb, d["q"], d["energy"], d["seeds"], d["color"], d["jump"], d["equilibrium"],
                                twin=True)
    return oracle_flash(mb, d["q"], d["energy"], d["seeds"], twin=True)


def check(d, states, ok, prof, analytic=False):
    ts, tok, tcnt = twin(d)
    assert np.array_equal(ok, tok)
    assert states.tobytes() == ts.tobytes()
    assert prof.phaseEvaluations == int(tcnt[:, 0].sum())
    assert prof.residualEvaluations == int(tcnt[:, 1].sum())
    assert prof.flashCandidates == int(tcnt[:, 2].sum())
    assert prof.analyticJacobians == int(tcnt[:, 5].sum())
    assert prof.finiteDifferenceJacobians == int(tcnt[:, 6].sum())
    # glibc's expm1, which every vapour-coordinate residual calls, is correctly
    # rounded on about 98 % of arguments (test_elementary.py), so more cells
    # than in the plain flash differ from the reference in a last bit (each
    # still within 1e-12 in every gated field)
    dist = assert_reference_parity(states, ok, d["out"], d["success"], analytic=analytic, max_not_bitwise=0.06)
    print(f"{len(ok)} cells, {int(ok.sum())} recovered, {dist['not_bitwise'].size} not bit for bit with the reference",
          flush=True)


@pytest.mark.parametrize("name", CAPILLARY)
def test_capillary_host_batch(name):
    d = load(name)
    states, ok, prof = host_batch(d, True)
    check(d, states, ok, prof)


@pytest.mark.parametrize("name", BOUNDARY)
def test_boundary_recovery_model(name):
    d = load(name)
    states, ok, prof = host_batch(d, False)
    check(d, states, ok, prof, analytic=name == "trace_analytic")


def test_capillary_resident_batch_equals_host_batch():
    """The same capillary cells from device pointers, in a scattered order."""
    d = load("trace_capillary")
    lib = rlib.load()
    mb = d["model"].tobytes()
    n = len(d["energy"])
    hs, hok, _ = host_batch(d, True)
    order = np.random.default_rng(3).permutation(n).astype(np.uint32)
    dev = Device()
    blank = np.zeros(n, abi.HEM_OUTPUT_DTYPE)
    dq, de, ds = dev.put(np.ascontiguousarray(d["q"])), dev.put(d["energy"]), dev.put(d["seeds"])
    dc, do, dout = dev.put(cap_records(d)), dev.put(order), dev.put(blank)
    err = C.create_string_buffer(1024)
    r = lib.reactive_gpu_hem_create_resident_v1(mb, len(mb), n, err, 1024)
    assert r, err.value
    prof = abi.HemProfile.fresh()
    assert lib.reactive_gpu_hem_run_resident_v1(r, dq, de, ds, dc, do, n, dout, C.byref(prof), err, 1024) == 0, \
        err.value
    out = dev.get(dout, blank)
    assert np.array_equal(out["success"], hok)
    good = hok == 1
    assert out["state"][good].tobytes() == hs[good].tobytes()
    dev.close()
    lib.reactive_gpu_hem_destroy_v1(r)


def test_boundary_model_rules():
    """The mixed arithmetic runs a boundaryRecovery model's flash in FP64 (the
    same results), and the resident step refuses such a model."""
    d = load("trace")
    lib = rlib.load()
    mb = d["model"].tobytes()
    n = len(d["energy"])
    ref, ref_ok, _ = host_batch(d, False)
    err = C.create_string_buffer(1024)
    h = lib.reactive_gpu_hem_create_v1(mb, len(mb), n, err, 1024)
    assert lib.reactive_amd_hem_set_arithmetic(h, MIXED) == 0
    states, ok = d["seeds"].copy(), np.zeros(n, np.int32)
    prof = abi.HemProfile.fresh()
    assert lib.reactive_gpu_hem_run_v1(h, rlib.ptr(np.ascontiguousarray(d["q"])), rlib.ptr(d["energy"]), n,
                                       rlib.ptr(states), rlib.ptr(ok), C.byref(prof), err, 1024) == 0
    assert np.array_equal(ok, ref_ok) and states.tobytes() == ref.tobytes()
    from helpers import config, cube_case
    case = cube_case(4, seed=2)
    cfg = config(len(case["q"]), len(case["faces"]), ns=2, nv=6)
    tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(case["volume"]), rlib.ptr(case["faces"]), None,
                                       None, None, None, err, 1024)
    assert tr, err.value
    assert not lib.reactive_amd_step_create(tr, h, err, 1024)
    assert b"boundary search" in err.value
    lib.reactive_transport_destroy(tr)
    lib.reactive_gpu_hem_destroy_v1(h)
