"""Transport parity on the GPU: the HIP library (backend 1) against the
oracle restatement of the reference's serial operators, bit for bit
(tools/validate_reactive_transport.py is the reference's own spec:
RHS, dt, conservation, stage algebra, token rules)."""
import ctypes as C

import numpy as np
import pytest

from helpers import OracleTransport, config, cube_case, jet_case, random_graph
from reactive_amd import abi, lib as rlib

pytestmark = pytest.mark.gpu


def create(lib, cfg, volume, faces, fixed_q=None, fixed_states=None, backend=1):
    err = C.create_string_buffer(1024)
    h = lib.reactive_transport_create(backend, C.byref(cfg), rlib.ptr(volume), rlib.ptr(faces), rlib.ptr(fixed_q),
                                      rlib.ptr(fixed_states), None, None, err, 1024)
    return h, err.value.decode()


def gpu_rhs(lib, h, q, st, nc, nv):
    rhs = np.zeros((nc, nv)); b = np.zeros(nv)
    rc = lib.reactive_transport_rhs(h, rlib.ptr(q), rlib.ptr(st), None, None, rlib.ptr(rhs), rlib.ptr(b))
    assert rc == 0, lib.reactive_transport_error(h)
    return rhs, b


@pytest.mark.parametrize("scheme", [0, 1], ids=["hll", "hllc"])
@pytest.mark.parametrize("n", [4, 7, 10, 34])
def test_cartesian_rhs_dt_stages_bitwise(n, scheme):
    lib = rlib.load()
    case = cube_case(n, seed=n)
    q, st = case["q"], case["tstates"]
    nc = len(q)
    cfg = config(nc, len(case["faces"]))
    h, err = create(lib, cfg, case["volume"], case["faces"])
    assert h, err
    assert lib.reactive_transport_is_cuda(h) == 1
    assert lib.reactive_amd_transport_set_flux(h, scheme) == 0
    ora = OracleTransport(cfg, case["volume"], case["faces"], scheme=scheme)
    r_gpu, b_gpu = gpu_rhs(lib, h, q, st, nc, 5)
    r_ora, b_ora = ora.rhs(q, st)
    assert np.array_equal(r_gpu, r_ora) and np.array_equal(b_gpu, b_ora)
    dt = C.c_double()
    assert lib.reactive_transport_stable_step(h, rlib.ptr(q), rlib.ptr(st), 0.4, 1.0, C.byref(dt)) == 0
    rc, dt_ora = ora.stable_step(q, st, 0.4, 1.0)
    assert rc == 0 and dt.value == dt_ora
    qg, qo = q.copy(), q.copy()
    for stage in (0, 1):
        bg = np.zeros(5)
        assert lib.reactive_transport_stage(h, rlib.ptr(qg), rlib.ptr(st), None, None, dt.value, stage,
                                            rlib.ptr(bg)) == 0, lib.reactive_transport_error(h)
        bo = ora.stage(qo, st, dt_ora, stage)
        assert np.array_equal(qg, qo), f"stage {stage}: max diff {np.abs(qg - qo).max()}"
        assert np.array_equal(bg, bo)
    # conservation: sum(rhs V) + boundary = 0 (validate_reactive_transport.py:206-208)
    total = (r_gpu * case["volume"][:, None]).sum(0) + b_gpu
    scale = np.abs(r_gpu * case["volume"][:, None]).sum(0) + np.abs(b_gpu) + 1e-300
    assert np.all(np.abs(total) <= 2e-11 * scale)
    lib.reactive_transport_destroy(h)
    ora.close()


@pytest.mark.parametrize("scheme", [0, 1], ids=["hll", "hllc"])
@pytest.mark.parametrize("slip", [(), (4, 5)], ids=["jets", "jets-slip"])
def test_cartesian_boundary_kinds_bitwise(slip, scheme):
    """Structured box with fixed-state jet inlets (kind 3) and slip walls
    (kind 1) on the fast path: rhs, dt and both stages bit for bit."""
    lib = rlib.load()
    case = jet_case(12, seed=21, slip_sides=slip)
    q, st = case["q"], case["tstates"]
    nc = len(q)
    fq, fst = case["fixed_q"], case["fixed_tstates"]
    cfg = config(nc, len(case["faces"]), fixed=len(fq))
    h, err = create(lib, cfg, case["volume"], case["faces"], fq, fst)
    assert h, err
    assert lib.reactive_amd_transport_set_flux(h, scheme) == 0
    ora = OracleTransport(cfg, case["volume"], case["faces"], fq, fst, scheme=scheme)
    r_gpu, b_gpu = gpu_rhs(lib, h, q, st, nc, 5)
    r_ora, b_ora = ora.rhs(q, st)
    assert np.array_equal(r_gpu, r_ora) and np.array_equal(b_gpu, b_ora)
    dt = C.c_double()
    assert lib.reactive_transport_stable_step(h, rlib.ptr(q), rlib.ptr(st), 0.4, 1.0, C.byref(dt)) == 0
    rc, dt_ora = ora.stable_step(q, st, 0.4, 1.0)
    assert rc == 0 and dt.value == dt_ora
    qg, qo = q.copy(), q.copy()
    for stage in (0, 1):
        bg = np.zeros(5)
        assert lib.reactive_transport_stage(h, rlib.ptr(qg), rlib.ptr(st), None, None, dt.value, stage,
                                            rlib.ptr(bg)) == 0, lib.reactive_transport_error(h)
        bo = ora.stage(qo, st, dt_ora, stage)
        assert np.array_equal(qg, qo), f"stage {stage}: max diff {np.abs(qg - qo).max()}"
        assert np.array_equal(bg, bo)
    lib.reactive_transport_destroy(h)
    ora.close()


@pytest.mark.parametrize("scheme", [0, 1], ids=["hll", "hllc"])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_generic_graph_all_face_kinds_bitwise(seed, scheme):
    lib = rlib.load()
    nc = 257
    faces, volume, q, st, fq, fst = random_graph(nc, seed)
    if scheme == 1:  # HLLC reads the recovered density, not sum(q): make them differ
        st = st.copy()
        st["rho"] *= 1 + 1e-3 * np.random.default_rng(seed).standard_normal(len(st))
    cfg = config(nc, len(faces), fixed=len(fq))
    h, err = create(lib, cfg, volume, faces, fq, fst)
    assert h, err
    assert lib.reactive_amd_transport_set_flux(h, scheme) == 0
    ora = OracleTransport(cfg, volume, faces, fq, fst, scheme=scheme)
    r_gpu, b_gpu = gpu_rhs(lib, h, q, st, nc, 5)
    r_ora, b_ora = ora.rhs(q, st)
    assert np.array_equal(r_gpu, r_ora) and np.array_equal(b_gpu, b_ora)
    dt = C.c_double()
    assert lib.reactive_transport_stable_step(h, rlib.ptr(q), rlib.ptr(st), 0.5, 1.0, C.byref(dt)) == 0
    assert dt.value == ora.stable_step(q, st, 0.5, 1.0)[1]
    qg, qo = q.copy(), q.copy()
    for stage in (0, 1):
        bg = np.zeros(5)
        assert lib.reactive_transport_stage(h, rlib.ptr(qg), rlib.ptr(st), None, None, dt.value, stage,
                                            rlib.ptr(bg)) == 0
        bo = ora.stage(qo, st, dt.value, stage)
        assert np.array_equal(qg, qo) and np.array_equal(bg, bo)
    lib.reactive_transport_destroy(h)
    ora.close()


def test_resident_tokens_and_rules():
    """Resident attempt protocol (reactiveTransport.cpp:1749-1789)."""
    lib = rlib.load()
    case = cube_case(5, seed=3)
    q, st = case["q"], case["tstates"]
    nc = len(q)
    cfg = config(nc, len(case["faces"]))
    opts = abi.OptionsV2(abiVersion=1, structBytes=C.sizeof(abi.OptionsV2), bridgeCells=64, recomputeGas=0,
                         physicalModelHash=b"a" * 64)
    err = C.create_string_buffer(1024)
    h = lib.reactive_transport_create_v2(1, C.byref(cfg), C.byref(opts), rlib.ptr(case["volume"]),
                                         rlib.ptr(case["faces"]), None, None, None, None, err, 1024)
    assert h, err.value
    ora = OracleTransport(cfg, case["volume"], case["faces"])
    dt = ora.stable_step(q, st, 0.4, 1.0)[1]
    b = np.zeros(5)
    assert lib.reactive_transport_begin_attempt(h, b"b" * 64, 1) == 1  # wrong hash
    assert lib.reactive_transport_begin_attempt(h, b"a" * 64, 1) == 0
    assert lib.reactive_transport_upload_conserved(h, rlib.ptr(q), 5) == 0
    T = abi.Token
    # stale stage token
    assert lib.reactive_transport_advance_resident_v2(h, T(1, 1, 5), T(1, 2, 6), rlib.ptr(st), None, None, None,
                                                      dt, rlib.ptr(b)) == 1
    assert lib.reactive_transport_upload_conserved(h, rlib.ptr(q), 7) == 0
    assert lib.reactive_transport_advance_resident_v2(h, T(1, 0, 7), T(1, 1, 8), rlib.ptr(st), None, None, None,
                                                      dt, rlib.ptr(b)) == 0, lib.reactive_transport_error(h)
    q1 = np.zeros_like(q)
    assert lib.reactive_transport_download_conserved(h, T(1, 1, 8), rlib.ptr(q1)) == 0
    qo = q.copy()
    ora.stage(qo, st, dt, 0)
    assert np.array_equal(q1, qo)
    assert lib.reactive_transport_advance_resident_v2(h, T(1, 1, 8), T(1, 2, 9), rlib.ptr(st), None, None, None,
                                                      dt, rlib.ptr(b)) == 0
    q2 = np.zeros_like(q)
    assert lib.reactive_transport_download_conserved(h, T(1, 2, 9), rlib.ptr(q2)) == 0
    ora.stage(qo, st, dt, 1)
    assert np.array_equal(q2, qo)
    assert lib.reactive_transport_end_attempt(h, 1, 1) == 0
    assert lib.reactive_transport_upload_conserved(h, rlib.ptr(q), 3) == 1  # versions must increase
    # a second attempt that stops after stage 0 cannot commit; it rolls back
    assert lib.reactive_transport_begin_attempt(h, b"a" * 64, 1) == 1  # attempt ids must increase
    assert lib.reactive_transport_begin_attempt(h, b"a" * 64, 2) == 0
    assert lib.reactive_transport_upload_conserved(h, rlib.ptr(q), 10) == 0
    assert lib.reactive_transport_advance_resident_v2(h, T(2, 0, 10), T(2, 1, 11), rlib.ptr(st), None, None, None,
                                                      dt, rlib.ptr(b)) == 0
    assert lib.reactive_transport_end_attempt(h, 2, 1) == 1
    assert b"incomplete" in lib.reactive_transport_error(h)
    assert lib.reactive_transport_end_attempt(h, 2, 0) == 0
    lib.reactive_transport_destroy(h)
    ora.close()


def test_replace_stage_state():
    """The post-RK1 state of an open attempt replaced between the stages
    (reactiveTransport.cpp:1636-1645; ReactiveFoam.C:1000): RK2 then combines
    the kept stage-0 input with the replacement, as the oracle does.  A
    refused replacement (no stage-1 state, a stale token or version, no
    array) fails like every failed call: the resident state is dropped
    (reactiveTransport.cpp:1481), so the attempt must roll back."""
    lib = rlib.load()
    case = cube_case(5, seed=4)
    q, st = case["q"], case["tstates"]
    cfg = config(len(q), len(case["faces"]))
    opts = abi.OptionsV2(abiVersion=1, structBytes=C.sizeof(abi.OptionsV2), bridgeCells=64, recomputeGas=0,
                         physicalModelHash=b"a" * 64)
    err = C.create_string_buffer(1024)
    h = lib.reactive_transport_create_v2(1, C.byref(cfg), C.byref(opts), rlib.ptr(case["volume"]),
                                         rlib.ptr(case["faces"]), None, None, None, None, err, 1024)
    assert h, err.value
    ora = OracleTransport(cfg, case["volume"], case["faces"])
    dt = ora.stable_step(q, st, 0.4, 1.0)[1]
    b = np.zeros(5)
    T = abi.Token
    try:
        assert lib.reactive_transport_begin_attempt(h, b"a" * 64, 1) == 0
        assert lib.reactive_transport_upload_conserved(h, rlib.ptr(q), 5) == 0
        assert lib.reactive_transport_advance_resident_v2(h, T(1, 0, 5), T(1, 1, 6), rlib.ptr(st), None, None, None,
                                                          dt, rlib.ptr(b)) == 0, lib.reactive_transport_error(h)
        qo = q.copy()
        ora.stage(qo, st, dt, 0)
        replaced = qo.copy()
        replaced[:, 4] *= 1.0 + 1e-3  # the driver's flash moved the energy
        assert lib.reactive_transport_replace_stage_state_v1(h, T(1, 1, 6), 7, rlib.ptr(replaced)) == 0, \
            lib.reactive_transport_error(h)
        got = np.zeros_like(q)
        assert lib.reactive_transport_download_conserved(h, T(1, 1, 7), rlib.ptr(got)) == 0
        assert np.array_equal(got, replaced)
        assert lib.reactive_transport_advance_resident_v2(h, T(1, 1, 7), T(1, 2, 8), rlib.ptr(st), None, None,
                                                          None, dt, rlib.ptr(b)) == 0, lib.reactive_transport_error(h)
        assert lib.reactive_transport_download_conserved(h, T(1, 2, 8), rlib.ptr(got)) == 0
        ora.stage(replaced, st, dt, 1)
        assert np.array_equal(got, replaced)
        assert lib.reactive_transport_end_attempt(h, 1, 1) == 0
        version = 10
        refusals = [lambda v: (T(2, 0, v), v + 1, rlib.ptr(q)),  # before stage 1
                    lambda v: (T(2, 1, v), v + 2, rlib.ptr(q)),  # stale content token
                    lambda v: (T(2, 1, v + 1), v + 1, rlib.ptr(q)),  # version not increasing
                    lambda v: (T(2, 1, v + 1), v + 2, None)]  # no array
        for attempt, refusal in enumerate(refusals, start=2):
            assert lib.reactive_transport_begin_attempt(h, b"a" * 64, attempt) == 0
            assert lib.reactive_transport_upload_conserved(h, rlib.ptr(q), version) == 0
            if attempt > 2:
                assert lib.reactive_transport_advance_resident_v2(
                    h, T(attempt, 0, version), T(attempt, 1, version + 1), rlib.ptr(st), None, None, None, dt,
                    rlib.ptr(b)) == 0, lib.reactive_transport_error(h)
            token, new_version, array = refusal(version)
            token = T(attempt, token.stageId, token.contentVersion)
            assert lib.reactive_transport_replace_stage_state_v1(h, token, new_version, array) == 1
            assert b"replacement" in lib.reactive_transport_error(h)
            # the resident state is gone: even the right replacement is refused now
            assert lib.reactive_transport_replace_stage_state_v1(h, T(attempt, 1, version + 1), version + 5,
                                                                 rlib.ptr(q)) == 1
            assert lib.reactive_transport_end_attempt(h, attempt, 1) == 1
            assert lib.reactive_transport_end_attempt(h, attempt, 0) == 0
            version += 10
    finally:
        lib.reactive_transport_destroy(h)
        ora.close()


def test_error_contracts():
    lib = rlib.load()
    case = cube_case(4, seed=4)
    q, st = case["q"], case["tstates"].copy()
    cfg = config(len(q), len(case["faces"]))
    h, err = create(lib, cfg, case["volume"], case["faces"], backend=0)
    assert not h and "backend=0" in err
    h, err = create(lib, cfg, case["volume"], case["faces"])
    b = np.zeros(5)
    qq = q.copy()
    assert lib.reactive_transport_stage(h, rlib.ptr(qq), rlib.ptr(st), None, None, 1e-7, 1, rlib.ptr(b)) == 1
    assert b"matching stage 0" in lib.reactive_transport_error(h)
    bad = st.copy()
    bad["sound"][3] = np.nan
    assert lib.reactive_transport_stage(h, rlib.ptr(qq), rlib.ptr(bad), None, None, 1e-7, 0, rlib.ptr(b)) == 1
    assert b"Invalid recovered transport state" in lib.reactive_transport_error(h)
    dt = C.c_double()
    assert lib.reactive_transport_stable_step(h, rlib.ptr(q), rlib.ptr(st), 0.6, 1.0, C.byref(dt)) == 1
    for position in (0, len(q) // 2, len(q) - 1):  # NaN velocity anywhere poisons the CFL minimum
        qn = q.copy()
        qn[position, 1] = np.nan
        assert lib.reactive_transport_stable_step(h, rlib.ptr(qn), rlib.ptr(st), 0.4, 1.0, C.byref(dt)) == 1
    assert lib.reactive_transport_set_wale_v1(h, None) == 1
    assert b"Null WALE options" in lib.reactive_transport_error(h)
    assert lib.reactive_transport_set_wale_scalars_v1(h, None) == 1
    assert b"not provided" in lib.reactive_transport_error(h)
    assert lib.reactive_transport_destroy(h) is None


def test_flux_scheme_contract():
    lib = rlib.load()
    case = cube_case(4, seed=2)
    cfg = config(len(case["q"]), len(case["faces"]))
    h, err = create(lib, cfg, case["volume"], case["faces"])
    assert h, err
    assert lib.reactive_amd_transport_set_flux(None, 1) == -1
    assert lib.reactive_amd_transport_set_flux(h, 7) == 1
    assert b"Unknown face flux scheme" in lib.reactive_transport_error(h)
    assert lib.reactive_amd_transport_set_flux(h, 1) == 0
    lib.reactive_transport_destroy(h)
