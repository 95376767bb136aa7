"""Device-resident SSPRK2 step (reactive_amd_step_run) against the same
step driven on the oracle: Flow::step without chemistry
(src/reactiveFoam/ReactiveFoam.C:961-1019).  Against the oracle's twin build
(the device's elementary functions) the whole step is compared bit for bit:
dt, q, every state field, failure counts.  Against the glibc build (the
reference itself) the step meets the parity gate: dt bit for bit, q within
1e-12, the second flash's states by helpers.assert_reference_parity, and
conserved totals within 1e-14."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from helpers import (OracleTransport, assert_reference_parity, binary_case, config, cube_case, jet_case,
                     oracle, oracle_flash)
from reactive_amd import abi, cases, lib as rlib

pytestmark = pytest.mark.gpu


THREADS = min(16, os.cpu_count() or 1)


def oracle_step(case, cfl, scheme=0, twin=True):
    """One oracle step; twin selects the flash build (device elementary
    functions) or glibc's (the reference)."""
    q0, s0, mb, ns = case["q"], case["states"], case["model"], case["ns"]
    fq = case.get("fixed_q")
    cfg = config(len(q0), len(case["faces"]), ns=ns, fixed=0 if fq is None else len(fq))
    ora = OracleTransport(cfg, case["volume"], case["faces"], fq, case.get("fixed_tstates"), scheme=scheme)
    rc, dt = ora.stable_step(q0, cases.transport_states(s0), cfl, 1.0)
    assert rc == 0
    q = q0.copy()
    bA = ora.stage(q, cases.transport_states(s0), dt, 0)
    s1, ok1, _ = oracle_flash(mb, q[:, :ns], cases.internal_energy(q, ns), s0, THREADS, twin)
    rc, recheck = ora.stable_step(q, cases.transport_states(s1), cfl, dt)
    bB = ora.stage(q, cases.transport_states(s1), dt, 1)
    s2, ok2, _ = oracle_flash(mb, q[:, :ns], cases.internal_energy(q, ns), s0, THREADS, twin)
    ora.close()
    return dict(dt=dt, recheck=recheck, q=q, s1=s1, s2=s2, ok=(ok1, ok2), bA=bA, bB=bB)


@pytest.mark.parametrize("n,scheme,kind", [(8, 0, "pure"), (12, 0, "pure"), (8, 1, "pure"), (8, 0, "binary"),
                                           (12, 0, "jets"), (12, 1, "jets-slip"), (10, 0, "pure-analytic"),
                                           (8, 0, "binary-analytic")],
                         ids=["8-hll", "12-hll", "8-hllc", "8-binary", "12-jets", "12-jets-slip-hllc",
                              "10-analytic-jacobian", "8-binary-analytic-jacobian"])
def test_resident_step_matches_oracle_step(n, scheme, kind):
    lib = rlib.load()
    if kind.startswith("binary"):
        case = binary_case(n, seed=5)
    elif kind.startswith("jets"):
        case = jet_case(n, seed=31, slip_sides=(4, 5) if kind == "jets-slip" else ())
    else:
        case = cube_case(n, seed=11 + n)
    analytic = kind.endswith("-analytic")
    if analytic:  # the reference's analytic-Jacobian flash (analyticJacobian = 1)
        from reactive_amd import model
        species = ("N2O", "N2") if case["ns"] == 2 else ("N2O",)
        case["model"] = model.model_bytes(model.build_model(gas_species=species, analytic_jacobian=1))
    cfl = 0.4
    ref = oracle_step(case, cfl, scheme)
    err = C.create_string_buffer(1024)
    mb = case["model"]
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(case["q"]), err, 1024)
    fq = case.get("fixed_q")
    cfg = config(len(case["q"]), len(case["faces"]), ns=case["ns"], fixed=0 if fq is None else len(fq))
    tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(case["volume"]), rlib.ptr(case["faces"]),
                                       rlib.ptr(fq), rlib.ptr(case.get("fixed_tstates")), None, None, err, 1024)
    assert lib.reactive_amd_transport_set_flux(tr, scheme) == 0
    step = lib.reactive_amd_step_create(tr, hem, err, 1024)
    assert step, err.value
    assert lib.reactive_amd_step_load(step, rlib.ptr(case["q"]), rlib.ptr(case["states"])) == 0
    rep = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, cfl, 1.0, 1, 0, C.byref(rep)) == 0, lib.reactive_amd_step_error(step)
    assert rep.fluxFailures == 0
    q = np.zeros_like(case["q"])
    s = np.zeros(len(q), dtype=abi.STATE_DTYPE)
    assert lib.reactive_amd_step_read(step, rlib.ptr(q), rlib.ptr(s)) == 0
    assert rep.dt == ref["dt"]  # same CFL arithmetic, bit for bit
    assert rep.flashFailures[0] == int((ref["ok"][0] == 0).sum())
    assert rep.flashFailures[1] == int((ref["ok"][1] == 0).sum())
    assert np.array_equal(q, ref["q"])
    assert s.tobytes() == ref["s2"].tobytes()
    # conservation totals: compensated device sum vs the exactly rounded sum of the same products
    vol = case["volume"][0]
    for k in range(q.shape[1]):
        exact = math.fsum(q[:, k] * vol)
        scale = math.fsum(np.abs(q[:, k] * vol))
        assert abs(rep.totals[k] - exact) <= 1e-14 * scale
    check_reference_step(case, cfl, scheme, rep, q, s, analytic)
    for k in range(q.shape[1]):
        assert rep.boundaryIntegral[k] == .5 * ref["dt"] * (ref["bA"][k] + ref["bB"][k]) or \
            abs(rep.boundaryIntegral[k] - .5 * ref["dt"] * (ref["bA"][k] + ref["bB"][k])) <= 1e-12 * (
                abs(ref["bA"][k]) + abs(ref["bB"][k])) * ref["dt"]
    lib.reactive_amd_step_destroy(step)
    lib.reactive_transport_destroy(tr)
    lib.reactive_gpu_hem_destroy_v1(hem)


def step_distance(q, r):
    """Largest difference of two conserved fields, relative per entry to the
    entry's magnitude or, for momenta near zero, to the largest momentum."""
    scale = np.maximum(np.abs(r), 1e-3 * np.abs(r).max(0))
    return float((np.abs(q - r) / scale).max())


def check_reference_step(case, cfl, scheme, rep, q, s, analytic=False):
    """The device step against the reference's arithmetic (glibc oracle step)."""
    ref = oracle_step(case, cfl, scheme, twin=False)
    assert rep.dt == ref["dt"]
    assert rep.flashFailures[1] == int((ref["ok"][1] == 0).sum())
    rq = step_distance(q, ref["q"])
    assert rq <= 1e-12, rq
    # a cell whose flash failed keeps its step-start state
    n = len(s)
    ok = (~(s.view(np.uint8).reshape(n, -1) == case["states"].view(np.uint8).reshape(n, -1)).all(1)).astype(np.int32)
    # (the second flash's inputs differ from the reference's in their last
    # bits wherever the first flash's p or c_f did, so cells need not be
    # bit for bit equal)
    d = assert_reference_parity(s, ok, ref["s2"], ref["ok"][1], max_not_bitwise=1.0, analytic=analytic)
    vol = case["volume"][0]
    for k in range(q.shape[1]):
        exact = math.fsum(ref["q"][:, k] * vol)
        scale = math.fsum(np.abs(ref["q"][:, k] * vol))
        assert abs(rep.totals[k] - exact) <= 1e-14 * scale
    print(f"step vs the reference: q max rel {rq:.1e}, {d['not_bitwise'].size} of {d['cells']} flash states "
          "not bit for bit,", {f: d[f][0] for f in ("p", "T", "soundFrozen", "soundEquilibrium", "alphaGas")})


def test_cube40_step_matches_oracle_and_reference():
    """The 40^3 pure-N2O box (64,000 cells, 80 mm, CFL 0.4, zero-gradient
    boundaries): one device step against the oracle step, bit for bit with the
    twin build and at the parity gate with the reference's arithmetic."""
    lib = rlib.load()
    case = cube_case(40, seed=2026, threads=THREADS)
    cfl = 0.4
    ref = oracle_step(case, cfl)
    err = C.create_string_buffer(1024)
    mb = case["model"]
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(case["q"]), err, 1024)
    cfg = config(len(case["q"]), len(case["faces"]))
    tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(case["volume"]), rlib.ptr(case["faces"]), None,
                                       None, None, None, err, 1024)
    step = lib.reactive_amd_step_create(tr, hem, err, 1024)
    assert step, err.value
    assert lib.reactive_amd_step_load(step, rlib.ptr(case["q"]), rlib.ptr(case["states"])) == 0
    rep = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, cfl, 1.0, 1, 0, C.byref(rep)) == 0, lib.reactive_amd_step_error(step)
    q = np.zeros_like(case["q"])
    s = np.zeros(len(q), dtype=abi.STATE_DTYPE)
    assert lib.reactive_amd_step_read(step, rlib.ptr(q), rlib.ptr(s)) == 0
    assert rep.dt == ref["dt"]
    assert np.array_equal(q, ref["q"])
    assert s.tobytes() == ref["s2"].tobytes()
    check_reference_step(case, cfl, 0, rep, q, s)
    lib.reactive_amd_step_destroy(step)
    lib.reactive_transport_destroy(tr)
    lib.reactive_gpu_hem_destroy_v1(hem)


def test_two_steps_cost_ordered():
    """The second step flashes its cells in the cost order recorded by the
    first; every cell must still be flashed exactly once, with the same result."""
    lib = rlib.load()
    n = 10
    case = cube_case(n, seed=29)
    cfl = 0.4
    first = oracle_step(case, cfl)
    second_case = dict(case, q=first["q"], states=first["s2"])  # failed cells kept their seeds
    second = oracle_step(second_case, cfl)
    err = C.create_string_buffer(1024)
    mb = case["model"]
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(case["q"]), err, 1024)
    cfg = config(len(case["q"]), len(case["faces"]))
    tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(case["volume"]), rlib.ptr(case["faces"]), None,
                                       None, None, None, err, 1024)
    step = lib.reactive_amd_step_create(tr, hem, err, 1024)
    assert step, err.value
    assert lib.reactive_amd_step_load(step, rlib.ptr(case["q"]), rlib.ptr(case["states"])) == 0
    rep = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, cfl, 1.0, 2, 0, C.byref(rep)) == 0, lib.reactive_amd_step_error(step)
    q = np.zeros_like(case["q"])
    s = np.zeros(len(q), dtype=abi.STATE_DTYPE)
    assert lib.reactive_amd_step_read(step, rlib.ptr(q), rlib.ptr(s)) == 0
    assert rep.dt == second["dt"]
    assert rep.flashFailures[1] == int((first["ok"][1] == 0).sum()) + int((second["ok"][1] == 0).sum())
    assert np.array_equal(q, second["q"])
    assert s.tobytes() == second["s2"].tobytes()
    lib.reactive_amd_step_destroy(step)
    lib.reactive_transport_destroy(tr)
    lib.reactive_gpu_hem_destroy_v1(hem)


@pytest.mark.parametrize("mode", [0, 1], ids=["fp64", "mixed"])
def test_graph_replays_match_stream_steps(mode):
    """Steps replayed from captured HIP graphs (REACTIVE_AMD_STEP_GRAPH) equal
    the same steps launched on the stream, bit for bit, replay after replay,
    with host reads (synchronisation and copies) between the replays."""
    lib = rlib.load()
    case = cube_case(10, seed=77)
    runs = {}
    for graph in (False, True):
        err = C.create_string_buffer(1024)
        mb = case["model"]
        hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(case["q"]), err, 1024)
        cfg = config(len(case["q"]), len(case["faces"]))
        tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(case["volume"]), rlib.ptr(case["faces"]),
                                           None, None, None, None, err, 1024)
        step = lib.reactive_amd_step_create(tr, hem, err, 1024)
        assert step, err.value
        assert lib.reactive_amd_step_load(step, rlib.ptr(case["q"]), rlib.ptr(case["states"])) == 0
        out = []
        for i in range(5):
            rep = abi.StepReport.fresh()
            flag = mode | (0x100 if graph else 0)
            assert lib.reactive_amd_step_run(step, 0.4, 1.0, 1 if i < 3 else 2, flag, C.byref(rep)) == 0, \
                lib.reactive_amd_step_error(step)
            q = np.zeros_like(case["q"])
            s = np.zeros(len(q), dtype=abi.STATE_DTYPE)
            assert lib.reactive_amd_step_read(step, rlib.ptr(q), rlib.ptr(s)) == 0
            out.append((q, s, rep.dt, tuple(rep.totals[:5])))
        runs[graph] = out
        lib.reactive_amd_step_destroy(step)
        lib.reactive_transport_destroy(tr)
        lib.reactive_gpu_hem_destroy_v1(hem)
    for (qe, se, dte, te), (qg, sg, dtg, tg) in zip(runs[False], runs[True]):
        assert np.array_equal(qe, qg)
        assert se.tobytes() == sg.tobytes()
        assert dte == dtg and te == tg


@pytest.mark.parametrize("mode", [0, 1], ids=["fp64", "mixed"])
def test_autotuned_shapes_leave_results_unchanged(mode):
    """REACTIVE_AMD_STEP_AUTOTUNE times the candidate launch shapes on the
    loaded state and keeps the fastest: the steps that follow equal untuned
    steps bit for bit, the trials leave the loaded state as it was, and the
    chosen shapes are candidates."""
    lib = rlib.load()
    case = cube_case(12, seed=83)
    runs = {}
    for tune in (False, True):
        err = C.create_string_buffer(1024)
        mb = case["model"]
        hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(case["q"]), err, 1024)
        cfg = config(len(case["q"]), len(case["faces"]))
        tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(case["volume"]), rlib.ptr(case["faces"]), None,
                                           None, None, None, err, 1024)
        step = lib.reactive_amd_step_create(tr, hem, err, 1024)
        assert step, err.value
        assert lib.reactive_amd_step_load(step, rlib.ptr(case["q"]), rlib.ptr(case["states"])) == 0
        rep = abi.StepReport.fresh()
        flags = mode | (0x200 if tune else 0)
        assert lib.reactive_amd_step_run(step, 0.4, 1.0, 2, flags, C.byref(rep)) == 0, lib.reactive_amd_step_error(step)
        q = np.zeros_like(case["q"])
        s = np.zeros(len(q), dtype=abi.STATE_DTYPE)
        assert lib.reactive_amd_step_read(step, rlib.ptr(q), rlib.ptr(s)) == 0
        runs[tune] = (q, s, rep.dt, (rep.tileDepth, rep.flashBlock, rep.stateBins))
        lib.reactive_amd_step_destroy(step)
        lib.reactive_transport_destroy(tr)
        lib.reactive_gpu_hem_destroy_v1(hem)
    print("\nuntuned shapes", runs[False][3], "tuned", runs[True][3])
    assert runs[False][3] == (4, 64, 64)
    depth, block, bins = runs[True][3]
    assert depth in (2, 4, 8) and block in (64, 128, 256) and bins in (32, 64, 128)
    assert runs[True][2] == runs[False][2]
    assert np.array_equal(runs[True][0], runs[False][0])
    assert runs[True][1].tobytes() == runs[False][1].tobytes()
