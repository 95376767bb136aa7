"""The oracle's viscous, conductive and WALE transport (oracle/oracle_transport.cpp)
against the reference's own serial operators (oracle/_ref/libref.so, written
into tests/golden/transport_viscous.npz by make_golden.py), bit for bit, and
against what the mathematics fixes: the Gauss gradient of a linear field, the
WALE eddy viscosity of pure shear (zero) and of solid-body rotation (closed
form), its homogeneity in the gradient, and conservation of the rates."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN, OracleTransport, config, oracle, p, primitives
from reactive_amd import abi, cases

FIXTURE = os.path.join(GOLDEN, "transport_viscous.npz")
MODELS = ["mu", "k", "mu_k", "wale", "wale_mu_k"]


def load(mesh, name):
    z = np.load(FIXTURE)
    base = {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(mesh + "__")}
    run = {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(f"{mesh}_{name}__")}
    return base, run


def oracle_for(base, run):
    fq = base["fixed_q"] if len(base["fixed_q"]) else None
    fst = base["fixed_states"] if len(base["fixed_states"]) else None
    cfg = config(len(base["q"]), len(base["faces"]), fixed=0 if fq is None else len(fq),
                 viscosity=float(run["viscosity"]), conductivity=float(run["conductivity"]))
    ora = OracleTransport(cfg, base["volume"], base["faces"], fq, fst)
    if float(run["cw"]) >= 0:
        ora.set_wale(float(run["cw"]))
    return ora


@pytest.mark.parametrize("name", MODELS)
@pytest.mark.parametrize("mesh", ["graph", "cube6"])
def test_oracle_matches_reference_viscous(mesh, name):
    base, run = load(mesh, name)
    q, ts = base["q"], base["states"]
    ora = oracle_for(base, run)
    if float(run["cw"]) >= 0:
        prim = primitives(q, 1)
        nut = ora.wale_primitives(prim, ts)
        assert np.array_equal(nut, run["nut"])
        assert (nut > 0).mean() > 0.5  # the closure is active on most cells
        rc, dtp = ora.stable_step_primitives(prim, ts, 0.4, 1.0)
        assert rc == 0 and dtp == float(run["dt_primitives"])
    rhs, b = ora.rhs(q, ts)
    assert np.array_equal(rhs, run["rhs"]) and np.array_equal(b, run["boundary"])
    rc, dt = ora.stable_step(q, ts, 0.4, 1.0)
    assert rc == 0 and dt == float(run["dt"])
    q1 = q.copy()
    b0 = ora.stage(q1, ts, dt, 0)
    assert np.array_equal(q1, run["q1"]) and np.array_equal(b0, run["b0"])
    q2 = q1.copy()
    b1 = ora.stage(q2, ts, dt, 1)
    assert np.array_equal(q2, run["q2"]) and np.array_equal(b1, run["b1"])
    ora.close()


@pytest.mark.parametrize("mesh", ["graph", "cube6"])
def test_viscous_terms_change_the_rates(mesh):
    """Each model moves the rates and the step away from the inviscid ones
    (the fixtures are not inviscid in disguise); conduction leaves the
    momentum rates alone and the viscous stress the species rates."""
    base, _ = load(mesh, "mu")
    q, ts = base["q"], base["states"]
    inviscid = oracle_for(base, dict(viscosity=0.0, conductivity=0.0, cw=-1.0))
    r0, _ = inviscid.rhs(q, ts)
    _, dt0 = inviscid.stable_step(q, ts, 0.4, 1.0)
    inviscid.close()
    for name in MODELS:
        _, run = load(mesh, name)
        r = run["rhs"]
        assert np.array_equal(r[:, 0], r0[:, 0])  # mass: no diffusion
        assert not np.array_equal(r[:, 4], r0[:, 4])
        if name == "k":
            assert np.array_equal(r[:, 1:4], r0[:, 1:4])
        else:
            assert not np.array_equal(r[:, 1:4], r0[:, 1:4])
        assert float(run["dt"]) <= dt0


@pytest.mark.parametrize("name", MODELS)
def test_viscous_rates_conserve(name):
    """sum_c rhs_c V_c + boundary = 0 per variable: every face flux enters
    its two cells with opposite signs."""
    base, run = load("graph", name)
    total = (run["rhs"] * base["volume"][:, None]).sum(0) + run["boundary"]
    scale = np.abs(run["rhs"] * base["volume"][:, None]).sum(0) + np.abs(run["boundary"])
    assert np.all(np.abs(total) <= 1e-12 * scale)


def test_gradient_of_a_linear_field_is_exact_inside():
    """u = A x on a uniform box: with ownerWeight 1/2 the face value is the
    field at the face centre, so the Gauss sum returns A on every interior
    cell (to rounding)."""
    n, h = 6, 1e-3
    faces, volume = cases.cartesian_faces(n, h)
    rng = np.random.default_rng(3)
    A = rng.normal(size=(3, 3)) * 100
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x = (np.stack([i.ravel(), j.ravel(), k.ravel()], 1) + 0.5) * h
    u = x @ A.T + rng.normal(size=3)
    rho = rng.uniform(10, 20, n ** 3)
    q = np.zeros((n ** 3, 5))
    q[:, 0] = rho
    q[:, 1:4] = rho[:, None] * u
    q[:, 4] = rho * 3e5
    st = np.zeros(n ** 3, dtype=abi.TSTATE_DTYPE)
    st["p"], st["T"], st["rho"], st["cv"], st["sound"], st["gasMass"] = 2e6, 300.0, rho, 900.0, 300.0, rho
    ora = OracleTransport(config(n ** 3, len(faces), viscosity=1e-3), volume, faces)
    ora.rhs(q, st)
    g = ora.gradient().reshape(-1, 3, 3)
    inside = ((i > 0) & (i < n - 1) & (j > 0) & (j < n - 1) & (k > 0) & (k < n - 1)).ravel()
    assert np.abs(g[inside] - A).max() <= 1e-9 * np.abs(A).max()
    ora.close()


def wale(g, delta=1e-3, cw=0.325):
    nut = C.c_double()
    ok = oracle().oracle_wale_eddy_viscosity(p(np.ascontiguousarray(g, dtype=float).ravel()), delta, cw,
                                             C.byref(nut))
    return ok, nut.value


def test_wale_closed_forms():
    # pure shear: g g = 0, so Sd = 0 and the WALE viscosity vanishes
    assert wale([[0, 250.0, 0], [0, 0, 0], [0, 0, 0]]) == (1, 0.0)
    # zero gradient, zero coefficient, zero width
    assert wale(np.zeros((3, 3))) == (1, 0.0)
    assert wale(np.eye(3), cw=0.0) == (1, 0.0)
    assert wale(np.eye(3), delta=0.0) == (1, 0.0)
    # solid-body rotation about z at rate w: S = 0, Sd:Sd = 2/3 w^4, so
    # nu_t = (Cw delta)^2 (Sd:Sd)^(1/4) = (Cw delta)^2 (2/3)^(1/4) |w|
    for w in (3.0, -1e4, 7e-200, 3e150):
        ok, nut = wale([[0, -w, 0], [w, 0, 0], [0, 0, 0]], delta=2e-3, cw=0.5)
        exact = (0.5 * 2e-3) ** 2 * (2 / 3) ** 0.25 * abs(w)
        assert ok == 1 and abs(nut - exact) <= 4e-16 * exact, (w, nut, exact)
    # homogeneous of degree one in g, scaled exactly by powers of two
    g = np.random.default_rng(5).normal(size=(3, 3)) * 40
    _, base = wale(g)
    for s in (2.0 ** -600, 0.5, 8.0, 2.0 ** 700):
        _, scaled = wale(g * s)
        assert scaled == base * s
    # failures: non-finite gradient, negative width or coefficient
    assert wale([[math.nan, 0, 0], [0, 0, 0], [0, 0, 0]])[0] == 0
    assert wale(np.eye(3), delta=-1.0)[0] == 0
    assert wale(np.eye(3), cw=-0.1)[0] == 0


def test_wale_install_rules():
    """installWale (reactiveTransport.cpp:636-652): once, before any upload."""
    base, run = load("graph", "wale")
    fq, fst = base["fixed_q"], base["fixed_states"]
    cfg = config(len(base["q"]), len(base["faces"]), fixed=len(fq))
    ora = OracleTransport(cfg, base["volume"], base["faces"], fq, fst)
    lib = ora.lib
    assert lib.oracle_transport_set_wale(ora.h, -1.0) == 1
    assert lib.oracle_transport_set_wale(ora.h, 0.325) == 0
    assert lib.oracle_transport_set_wale(ora.h, 0.325) == 1
    ora.close()
    ora = OracleTransport(cfg, base["volume"], base["faces"], fq, fst)
    ora.rhs(base["q"], base["states"])
    assert ora.lib.oracle_transport_set_wale(ora.h, 0.325) == 1
    ora.close()
