"""The resolved-interface capillary model on the GPU (transport.hip): the HIP
library (backend 1) against the reference's own serial operators with the
model installed (tests/golden/transport_capillary.npz, written from
oracle/_ref/libref.so by tests/golden/make_golden.py), bit for bit: the
colour geometry (surface energy, curvature, normals), rhs and boundary rates,
the CFL step with the capillary wave step, both RK stages and the capillary
profile counters; on the oblique graph with every face kind and on a 6^3
N2O box, at sigma = 0.02 N/m, at 1e6 N/m (capillary-dominated: the
capillary wave step sets dt) and with viscosity and conduction.  Then the
rules around the model (reactiveTransport.cpp:859-962, 1381, 1405-1408)."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import config, random_graph
from reactive_amd import abi, lib as rlib

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "transport_capillary.npz")
MODELS = ("s002", "s1e6", "s002_mu_k")
PROFILE_FIELDS = [f for f, _ in abi.CapillaryProfile._fields_[2:]]


def load(mesh, name):
    z = np.load(FIXTURE)
    base = {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(mesh + "__")}
    run = {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(f"{mesh}_{name}__")}
    return base, run


def create(lib, base, mu=0.0, k=0.0, nv=6):
    fq = base["fixed_q"] if len(base["fixed_q"]) else None
    fst = base["fixed_states"] if len(base["fixed_states"]) else None
    if fq is not None and fq.shape[1] != nv:
        fq = np.ascontiguousarray(fq[:, :nv])
    cfg = config(len(base["q"]), len(base["faces"]), nv=nv, fixed=0 if fq is None else len(fq), viscosity=mu,
                 conductivity=k)
    err = C.create_string_buffer(1024)
    h = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(base["volume"]), rlib.ptr(base["faces"]),
                                      rlib.ptr(fq), rlib.ptr(fst), None, None, err, 1024)
    assert h, err.value
    return h


def install(lib, h, sigma, ccfl=0.5, eps=1e-6, species=0):
    opt = abi.CapillaryOptions.make(sigma, ccfl, eps, species)
    return lib.reactive_transport_set_capillary_v1(h, C.byref(opt))


def geometry(lib, h, base, normals=True):
    n = len(base["q"])
    energy, curvature, normal = np.zeros(n), np.zeros(n), np.zeros((n, 3))
    fixed = base["fixed_colour"] if len(base["fixed_colour"]) else None
    rc = lib.reactive_transport_capillary_geometry_v1(h, rlib.ptr(base["colour"]), rlib.ptr(fixed),
                                                      rlib.ptr(energy), rlib.ptr(curvature),
                                                      rlib.ptr(normal) if normals else None)
    return rc, energy, curvature, normal


def profile(lib, h):
    prof = abi.CapillaryProfile.fresh()
    assert lib.reactive_transport_capillary_profile_v1(h, C.byref(prof)) == 0, lib.reactive_transport_error(h)
    return np.array([getattr(prof, f) for f in PROFILE_FIELDS], np.uint64)


def run_all(lib, h, q, st, nv=6, cfl=0.4):
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
    return dict(rhs=rhs, boundary=b, dt=dt.value, q1=q1, b0=b0, q2=q2, b1=b1)


@pytest.mark.parametrize("name", MODELS)
@pytest.mark.parametrize("mesh", ["graph", "cube6"])
def test_capillary_transport_matches_reference_bitwise(mesh, name):
    lib = rlib.load()
    base, ref = load(mesh, name)
    sigma, ccfl, eps, mu, k = (float(x) for x in ref["model"])
    print(f"capillary transport {mesh} {name}: sigma {sigma}, viscosity {mu}, conductivity {k}", flush=True)
    h = create(lib, base, mu, k)
    try:
        assert install(lib, h, sigma, ccfl, eps) == 0, lib.reactive_transport_error(h)
        rc, energy, curvature, normal = geometry(lib, h, base)
        assert rc == 0, lib.reactive_transport_error(h)
        for key, got in (("energy", energy), ("curvature", curvature), ("normal", normal)):
            assert np.array_equal(got, ref[key]), (key, np.abs(got - ref[key]).max())
        got = run_all(lib, h, base["q"].copy(), base["states"])
        for key in ("rhs", "boundary", "dt", "q1", "b0", "q2", "b1"):
            assert np.array_equal(got[key], ref[key]), (key, np.max(np.abs(np.asarray(got[key]) - ref[key])))
        assert np.array_equal(profile(lib, h), ref["profile"]), (profile(lib, h), ref["profile"])
    finally:
        lib.reactive_transport_destroy(h)


def test_capillary_geometry_reuses_an_identical_request():
    """A geometry request with the same colour bytes and no normals returns
    the kept outputs without device work (reactiveTransport.cpp:900-904): the
    profile does not move; a changed colour rebuilds."""
    lib = rlib.load()
    base, ref = load("cube6", "s002")
    h = create(lib, base)
    try:
        assert install(lib, h, 0.02) == 0, lib.reactive_transport_error(h)
        rc, energy, curvature, _ = geometry(lib, h, base, normals=False)
        assert rc == 0, lib.reactive_transport_error(h)
        first = profile(lib, h)
        rc, energy2, curvature2, _ = geometry(lib, h, base, normals=False)
        assert rc == 0 and np.array_equal(profile(lib, h), first)
        assert np.array_equal(energy2, energy) and np.array_equal(curvature2, curvature)
        assert np.array_equal(energy, ref["energy"]) and np.array_equal(curvature, ref["curvature"])
        changed = dict(base, colour=np.clip(base["colour"] * 0.5, 0, 1))
        rc, energy3, _, _ = geometry(lib, h, changed, normals=False)
        assert rc == 0 and profile(lib, h)[0] == first[0] + 1
        assert not np.array_equal(energy3, energy)
    finally:
        lib.reactive_transport_destroy(h)


def test_capillary_rules():
    """The model's contracts: options, the variable layout, install once
    before stepping, geometry before a stage, colours in [0, 1], the liquid
    inventory bounds on every face, and the profile ABI."""
    lib = rlib.load()
    base, _ = load("graph", "s002")
    h = create(lib, base)
    try:
        bad = abi.CapillaryOptions.make(0.02)
        bad.sigma = 0.0
        assert lib.reactive_transport_set_capillary_v1(h, C.byref(bad)) != 0
        assert b"Invalid capillary options" in lib.reactive_transport_error(h)
        assert install(lib, h, 0.02, ccfl=1.5) != 0
        assert install(lib, h, 0.02, species=1) != 0
        assert b"liquid inventory" in lib.reactive_transport_error(h)
        assert lib.reactive_transport_set_capillary_v1(h, None) != 0
        prof = abi.CapillaryProfile(abiVersion=2, structBytes=C.sizeof(abi.CapillaryProfile))
        assert lib.reactive_transport_capillary_profile_v1(h, C.byref(prof)) != 0
        # no model yet: geometry is refused
        rc, *_ = geometry(lib, h, base)
        assert rc != 0 and b"Missing capillary model" in lib.reactive_transport_error(h)
        assert install(lib, h, 0.02) == 0, lib.reactive_transport_error(h)
        assert install(lib, h, 0.02) != 0
        assert b"installed once" in lib.reactive_transport_error(h)
        q, st = base["q"].copy(), base["states"]
        rhs, b = np.zeros_like(q), np.zeros(6)
        assert lib.reactive_transport_rhs(h, rlib.ptr(q), rlib.ptr(st), None, None, rlib.ptr(rhs), rlib.ptr(b)) != 0
        assert b"Capillary geometry must be uploaded" in lib.reactive_transport_error(h)
        outside = dict(base, colour=base["colour"].copy())
        outside["colour"][5] = 1.5
        rc, *_ = geometry(lib, h, outside)
        assert rc != 0 and b"Invalid material color" in lib.reactive_transport_error(h)
        rc, *_ = geometry(lib, h, base)
        assert rc == 0, lib.reactive_transport_error(h)
        over = q.copy()
        over[17, 5] = 1.5 * over[17, 0]  # liquid inventory above the species' own mass
        assert lib.reactive_transport_rhs(h, rlib.ptr(over), rlib.ptr(st), None, None, rlib.ptr(rhs),
                                          rlib.ptr(b)) != 0
        assert b"Invalid capillary face flux" in lib.reactive_transport_error(h)
        assert lib.reactive_transport_rhs(h, rlib.ptr(q), rlib.ptr(st), None, None, rlib.ptr(rhs), rlib.ptr(b)) == 0, \
            lib.reactive_transport_error(h)
    finally:
        lib.reactive_transport_destroy(h)
    # five variables: no liquid inventory to map
    fg, vg, qg, tg, fqg, ftg = random_graph(40, 3)
    cfg = config(len(qg), len(fg), fixed=len(fqg))
    err = C.create_string_buffer(1024)
    h = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(vg), rlib.ptr(fg), rlib.ptr(fqg), rlib.ptr(ftg),
                                      None, None, err, 1024)
    assert h, err.value
    try:
        assert install(lib, h, 0.02) != 0
        assert b"liquid inventory" in lib.reactive_transport_error(h)
    finally:
        lib.reactive_transport_destroy(h)
