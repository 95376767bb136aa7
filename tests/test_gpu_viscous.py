"""Viscous stress, conduction and the WALE closure on the GPU (transport.hip):
the HIP library (backend 1) against the reference's own serial operators
(tests/golden/transport_viscous.npz, written from oracle/_ref/libref.so) and
against the oracle, bit for bit: rhs and boundary rates, CFL step (with the
diffusion number), both RK stages, the WALE eddy viscosities and the CFL step
from primitives; then larger boxes, every face kind, the HLLC option, the
resident stage protocol and the rules around the closure."""
import ctypes as C

import numpy as np
import pytest

from helpers import OracleTransport, config, cube_case, jet_case, primitives, random_graph
from reactive_amd import abi, lib as rlib
from test_oracle_viscous import MODELS, load

pytestmark = pytest.mark.gpu


def create(lib, cfg, volume, faces, fixed_q=None, fixed_states=None):
    err = C.create_string_buffer(1024)
    h = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(volume), rlib.ptr(faces), rlib.ptr(fixed_q),
                                      rlib.ptr(fixed_states), None, None, err, 1024)
    assert h, err.value
    return h


def set_wale(lib, h, cw):
    opt = abi.WaleOptions.make(cw)
    assert lib.reactive_transport_set_wale_v1(h, C.byref(opt)) == 0, lib.reactive_transport_error(h)


def run_all(lib, h, q, st, nv, wale, cfl=0.4):
    """nut and dt from primitives (WALE only), rhs, dt, stage 0, stage 1."""
    out = {}
    if wale:
        prim = primitives(q, nv - 4)
        nut = np.zeros(len(q))
        assert lib.reactive_transport_wale_primitives_v1(h, rlib.ptr(prim), rlib.ptr(st), rlib.ptr(nut)) == 0, \
            lib.reactive_transport_error(h)
        dtp = C.c_double()
        assert lib.reactive_transport_stable_step_primitives(h, rlib.ptr(prim), rlib.ptr(st), cfl, 1.0,
                                                             C.byref(dtp)) == 0, lib.reactive_transport_error(h)
        out.update(nut=nut, dt_primitives=dtp.value)
    rhs = np.zeros_like(q)
    b = np.zeros(nv)
    assert lib.reactive_transport_rhs(h, rlib.ptr(q), rlib.ptr(st), None, None, rlib.ptr(rhs), rlib.ptr(b)) == 0, \
        lib.reactive_transport_error(h)
    dt = C.c_double()
    assert lib.reactive_transport_stable_step(h, rlib.ptr(q), rlib.ptr(st), cfl, 1.0, C.byref(dt)) == 0, \
        lib.reactive_transport_error(h)
    q1 = q.copy()
    b0 = np.zeros(nv)
    assert lib.reactive_transport_stage(h, rlib.ptr(q1), rlib.ptr(st), None, None, dt.value, 0, rlib.ptr(b0)) == 0, \
        lib.reactive_transport_error(h)
    q2 = q1.copy()
    b1 = np.zeros(nv)
    assert lib.reactive_transport_stage(h, rlib.ptr(q2), rlib.ptr(st), None, None, dt.value, 1, rlib.ptr(b1)) == 0, \
        lib.reactive_transport_error(h)
    out.update(rhs=rhs, boundary=b, dt=dt.value, q1=q1, b0=b0, q2=q2, b1=b1)
    return out


def oracle_all(ora, q, st, nv, wale, cfl=0.4):
    out = {}
    if wale:
        prim = primitives(q, nv - 4)
        out["nut"] = ora.wale_primitives(prim, st)
        rc, out["dt_primitives"] = ora.stable_step_primitives(prim, st, cfl, 1.0)
        assert rc == 0
    out["rhs"], out["boundary"] = ora.rhs(q, st)
    rc, out["dt"] = ora.stable_step(q, st, cfl, 1.0)
    assert rc == 0
    q1 = q.copy()
    out["b0"] = ora.stage(q1, st, out["dt"], 0)
    q2 = q1.copy()
    out["b1"] = ora.stage(q2, st, out["dt"], 1)
    out.update(q1=q1, q2=q2)
    return out


def assert_same(got, want, keys):
    for k in keys:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        assert np.array_equal(g, w), f"{k}: max diff {np.abs(g - w).max()}"


@pytest.mark.parametrize("name", MODELS)
@pytest.mark.parametrize("mesh", ["graph", "cube6"])
def test_viscous_matches_reference_bitwise(mesh, name):
    lib = rlib.load()
    base, run = load(mesh, name)
    q, ts = base["q"], base["states"]
    fq = base["fixed_q"] if len(base["fixed_q"]) else None
    fst = base["fixed_states"] if len(base["fixed_states"]) else None
    cfg = config(len(q), len(base["faces"]), fixed=0 if fq is None else len(fq),
                 viscosity=float(run["viscosity"]), conductivity=float(run["conductivity"]))
    h = create(lib, cfg, base["volume"], base["faces"], fq, fst)
    wale = float(run["cw"]) >= 0
    if wale:
        set_wale(lib, h, float(run["cw"]))
    got = run_all(lib, h, q, ts, 5, wale)
    keys = ["rhs", "boundary", "dt", "q1", "b0", "q2", "b1"] + (["nut", "dt_primitives"] if wale else [])
    assert_same(got, run, keys)
    if wale:
        prof = abi.WaleProfile.fresh()
        assert lib.reactive_transport_wale_profile_v1(h, C.byref(prof)) == 0
        # wale_primitives, stable_step_primitives, rhs, stable_step, two stages
        assert prof.gradientBuilds == prof.viscosityBuilds == 6 and prof.cellsEvaluated == 6 * len(q)
        assert prof.workspaceBytes > 0
    lib.reactive_transport_destroy(h)


@pytest.mark.parametrize("scheme", [0, 1], ids=["hll", "hllc"])
@pytest.mark.parametrize("model", [(1e-4, 0.02, None), (5e-5, 0.01, 0.325), (0.0, 0.0, 0.5)],
                         ids=["mu_k", "wale_mu_k", "wale"])
@pytest.mark.parametrize("n", [13, 34])
def test_viscous_box_against_oracle(n, model, scheme):
    """Structured boxes (the generic kernels carry every viscous box) with
    flash-consistent N2O states, against the oracle."""
    lib = rlib.load()
    mu, k, cw = model
    case = cube_case(n, seed=40 + n)
    q, st = case["q"], case["tstates"]
    cfg = config(len(q), len(case["faces"]), viscosity=mu, conductivity=k)
    h = create(lib, cfg, case["volume"], case["faces"])
    assert lib.reactive_amd_transport_set_flux(h, scheme) == 0
    ora = OracleTransport(cfg, case["volume"], case["faces"], scheme=scheme)
    if cw is not None:
        set_wale(lib, h, cw)
        ora.set_wale(cw)
    assert_same(run_all(lib, h, q, st, 5, cw is not None), oracle_all(ora, q, st, 5, cw is not None),
                ["rhs", "boundary", "dt", "q1", "b0", "q2", "b1"] + (["nut", "dt_primitives"] if cw else []))
    lib.reactive_transport_destroy(h)
    ora.close()


@pytest.mark.parametrize("slip", [(), (4, 5)], ids=["jets", "jets-slip"])
def test_viscous_jets_with_walls_against_oracle(slip):
    """Fixed-state inlets and slip walls: the two-point gradient across a
    wall uses twice the distance and the wall keeps only the normal stress."""
    lib = rlib.load()
    case = jet_case(12, seed=23, slip_sides=slip)
    q, st = case["q"], case["tstates"]
    fq, fst = case["fixed_q"], case["fixed_tstates"]
    cfg = config(len(q), len(case["faces"]), fixed=len(fq), viscosity=2e-4, conductivity=0.05)
    h = create(lib, cfg, case["volume"], case["faces"], fq, fst)
    ora = OracleTransport(cfg, case["volume"], case["faces"], fq, fst)
    set_wale(lib, h, 0.325)
    ora.set_wale(0.325)
    assert_same(run_all(lib, h, q, st, 5, True), oracle_all(ora, q, st, 5, True),
                ["rhs", "boundary", "dt", "q1", "b0", "q2", "b1", "nut", "dt_primitives"])
    lib.reactive_transport_destroy(h)
    ora.close()


@pytest.mark.parametrize("seed", [11, 12])
def test_viscous_graph_resident_stages(seed):
    """The resident stage entry (upload_conserved + stage_resident) of a
    viscous, conductive WALE graph: both stages against the oracle."""
    lib = rlib.load()
    faces, volume, q, st, fq, fst = random_graph(301, seed)
    cfg = config(len(q), len(faces), fixed=len(fq), viscosity=0.02, conductivity=25.0)
    h = create(lib, cfg, volume, faces, fq, fst)
    ora = OracleTransport(cfg, volume, faces, fq, fst)
    set_wale(lib, h, 0.4)
    ora.set_wale(0.4)
    rc, dt = ora.stable_step(q, st, 0.4, 1.0)
    assert rc == 0
    assert lib.reactive_transport_upload_conserved(h, rlib.ptr(q), 1) == 0
    qo = q.copy()
    out = np.zeros_like(q)
    for stage in (0, 1):
        b = np.zeros(5)
        assert lib.reactive_transport_stage_resident(h, rlib.ptr(st), None, None, dt, stage, 1 + stage, 2 + stage,
                                                     rlib.ptr(out), rlib.ptr(b)) == 0, lib.reactive_transport_error(h)
        bo = ora.stage(qo, st, dt, stage)
        assert np.array_equal(out, qo) and np.array_equal(b, bo)
    lib.reactive_transport_destroy(h)
    ora.close()


def test_viscous_rules():
    lib = rlib.load()
    case = cube_case(6, seed=8)
    q, st = case["q"], case["tstates"]
    err = C.create_string_buffer(1024)
    # species diffusion stays out of this build
    cfg = config(len(q), len(case["faces"]))
    cfg.diffusivity = 1e-5
    assert not lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(case["volume"]), rlib.ptr(case["faces"]),
                                             None, None, rlib.ptr(np.zeros(1)), rlib.ptr(np.zeros(1)), err, 1024)
    assert b"diffusivity" in err.value
    cfg = config(len(q), len(case["faces"]), viscosity=-1.0)
    assert not lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(case["volume"]), rlib.ptr(case["faces"]),
                                             None, None, None, None, err, 1024)
    assert b"Invalid transport coefficients" in err.value
    # the resident step takes the inviscid flux only
    cfg = config(len(q), len(case["faces"]), viscosity=1e-4)
    h = create(lib, cfg, case["volume"], case["faces"])
    hem = lib.reactive_gpu_hem_create_v1(case["model"], len(case["model"]), len(q), err, 1024)
    assert hem, err.value
    assert not lib.reactive_amd_step_create(h, hem, err, 1024)
    assert b"inviscid" in err.value
    lib.reactive_transport_destroy(h)
    # WALE: once, before any upload, with a valid coefficient, not on a box a step uses
    cfg = config(len(q), len(case["faces"]))
    h = create(lib, cfg, case["volume"], case["faces"])
    bad = abi.WaleOptions.make(-0.1)
    assert lib.reactive_transport_set_wale_v1(h, C.byref(bad)) == 1
    assert b"Invalid WALE coefficient" in lib.reactive_transport_error(h)
    nut = np.zeros(len(q))
    prim = primitives(q, 1)
    assert lib.reactive_transport_wale_primitives_v1(h, rlib.ptr(prim), rlib.ptr(st), rlib.ptr(nut)) == 1
    assert b"WALE not selected" in lib.reactive_transport_error(h)
    set_wale(lib, h, 0.325)
    opt = abi.WaleOptions.make(0.325)
    assert lib.reactive_transport_set_wale_v1(h, C.byref(opt)) == 1
    assert b"immutable" in lib.reactive_transport_error(h)
    assert not lib.reactive_amd_step_create(h, hem, err, 1024)
    lib.reactive_transport_destroy(h)
    h = create(lib, cfg, case["volume"], case["faces"])
    rhs, b = np.zeros_like(q), np.zeros(5)
    assert lib.reactive_transport_rhs(h, rlib.ptr(q), rlib.ptr(st), None, None, rlib.ptr(rhs), rlib.ptr(b)) == 0
    assert lib.reactive_transport_set_wale_v1(h, C.byref(opt)) == 1
    lib.reactive_transport_destroy(h)
    h = create(lib, cfg, case["volume"], case["faces"])
    step = lib.reactive_amd_step_create(h, hem, err, 1024)
    assert step, err.value
    assert lib.reactive_transport_set_wale_v1(h, C.byref(opt)) == 1
    assert b"resident step" in lib.reactive_transport_error(h)
    lib.reactive_amd_step_destroy(step)
    lib.reactive_transport_destroy(h)
    lib.reactive_gpu_hem_destroy_v1(hem)
