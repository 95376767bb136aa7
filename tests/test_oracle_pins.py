"""Pins of the oracle that do not go through the reference: closed forms,
invariants and library routines, chosen so that a dropped term, a wrong sign
or index or a transposed operand fails one of them."""
import ctypes as C
import math

import numpy as np
import pytest

from helpers import OracleTransport, config, oracle, oracle_flash
from reactive_amd import abi, cases, mechanism, model

R = mechanism.GAS_CONSTANT


def mbuf():
    mb = model.model_bytes(model.build_model())
    return C.create_string_buffer(mb, len(mb)), mb


def test_lu_matches_numpy():
    rng = np.random.default_rng(0)
    lib = oracle()
    for n in (2, 3, 4):
        for _ in range(200):
            a = rng.normal(size=(n, n))
            b = rng.normal(size=n)
            x = np.zeros(n)
            assert lib.oracle_solve(a.ctypes.data, b.ctypes.data, n, x.ctypes.data) == 1
            assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)
    singular = np.array([[1.0, 2.0], [2.0, 4.0]])
    assert lib.oracle_solve(singular.ctypes.data, np.ones(2).ctypes.data, 2, np.zeros(2).ctypes.data) == 0


def test_nasa9_matches_polynomial():
    buf, _ = mbuf()
    d = mechanism.SPECIES["N2O"]
    out = np.zeros(3)
    for T in (150.0, 182.34, 190.0, 200.0, 650.0, 1420.0, 2500.0):
        oracle().oracle_nasa(buf, 0, 0, T, out.ctypes.data)
        r = np.searchsorted(d["ranges"][1:-1], T, side="right")  # NASA9: upper region at a shared bound
        a = d["nasa9"][r]
        cp = a[0] / T**2 + a[1] / T + a[2] + a[3] * T + a[4] * T**2 + a[5] * T**3 + a[6] * T**4
        h = -a[0] / T**2 + a[1] * math.log(T) / T + a[2] + a[3] * T / 2 + a[4] * T**2 / 3 + a[5] * T**3 / 4 \
            + a[6] * T**4 / 5 + a[7] / T
        s = -a[0] / T**2 / 2 - a[1] / T + a[2] * math.log(T) + a[3] * T + a[4] * T**2 / 2 + a[5] * T**3 / 3 \
            + a[6] * T**4 / 4 + a[8]
        assert out == pytest.approx([cp, h, s], rel=1e-13)


def test_pr_state_matches_textbook():
    """p(T, V), e(T, V) and the cubic: Peng-Robinson from its definition."""
    buf, _ = mbuf()
    fluid = cases.PureFluid("N2O")
    Y = np.array([1.0] + [0.0] * 15)
    out = np.zeros(32)
    for T, rho in ((300.0, 5.0), (250.0, 40.0), (200.0, 1250.0), (320.0, 300.0)):
        assert oracle().oracle_pr_trho(buf, 0, T, rho, Y.ctypes.data, out.ctypes.data) == 0
        V = fluid.W / rho
        assert out[1] == pytest.approx(fluid.pressure(T, V), rel=1e-12)
        assert out[5] == pytest.approx(fluid.energy(T, rho), rel=1e-10)
        # the state's V solves the cubic at its own pressure: p(T, V) = p
        roots = np.zeros(7)
        x = np.array([1.0])
        assert oracle().oracle_pr_roots(buf, 0, T, out[1], x.ctypes.data, roots.ctypes.data) == 0
        found = roots[:int(roots[3])]
        assert np.min(np.abs(found - V)) <= 1e-9 * V


def test_cubic_root_count_by_brute_force():
    buf, _ = mbuf()
    fluid = cases.PureFluid("N2O")
    x = np.array([1.0])
    roots = np.zeros(7)
    for T, p in ((200.0, 1e5), (250.0, 2.0e6), (280.0, 3.0e6), (300.0, 2e5), (350.0, 1e7)):
        assert oracle().oracle_pr_roots(buf, 0, T, p, x.ctypes.data, roots.ctypes.data) == 0
        V = np.geomspace(fluid.b * (1 + 1e-9), 1e3, 200001)
        f = fluid.pressure(T, V) - p
        crossings = int(np.count_nonzero(np.sign(f[1:]) != np.sign(f[:-1])))
        assert int(roots[3]) == crossings, (T, p)


def test_coexistence_equal_potentials_and_two_phase_flash():
    """At the saturation pressure the two roots have equal chemical potential
    (validate_n2o_supercooled.py:27-40); a lever-rule state flashes back to T."""
    buf, mb = mbuf()
    fluid = cases.PureFluid("N2O")
    pure = np.array([1.0] + [0.0] * 15)
    for T in (160.0, 170.0, 182.0, 190.0, 250.0):
        psat, vl, vv = fluid.saturation(T)
        gas = np.zeros(32)
        liq = np.zeros(32)
        assert oracle().oracle_pr_tp(buf, 0, T, psat, pure.ctypes.data, 1, gas.ctypes.data) == 0
        assert oracle().oracle_pr_tp(buf, 1, T, psat, pure.ctypes.data, 2, liq.ctypes.data) == 0
        assert abs(gas[15] - liq[15]) <= 1e-7 * R * T  # mu in J/kmol
        rho = fluid.W / (0.5 * (vl + vv))
        e, p, liquid, dome = cases.equilibrium_states(np.array([T]), np.array([rho]), fluid)
        seeds = np.zeros(1, dtype=abi.STATE_DTYPE)
        seeds["p"] = p; seeds["T"] = T; seeds["liquidMass"][:, 0] = liquid; seeds["activeLiquids"] = 1
        out, ok, _ = oracle_flash(mb, np.array([[rho]]), rho * e, seeds)
        assert ok[0] == 1 and out["liquidMass"][0, 0] > 0 and out["gasMass"][0] > 0
        assert out["T"][0] == pytest.approx(T, abs=2e-4)  # saturation tabulated on a grid


def test_flash_invariants_and_gas_round_trip():
    """Converged states satisfy the residual tolerances and c_eq <= c_f (1 + 2e-3);
    a single-phase gas state e(rho, T) flashes back to T."""
    _, mb = mbuf()
    faces, volume, q, seeds, _ = cases.pure_n2o_cube(6, seed=21)
    e = cases.internal_energy(q, 1)
    out, ok, _ = oracle_flash(mb, q[:, :1], e, seeds)
    good = ok == 1
    assert good.mean() > 0.98
    assert (out["volumeResidual"][good] <= 1e-9).all() and (out["energyResidual"][good] <= 1e-9).all()
    assert (out["chemicalResidual"][good] <= 1e-7).all()
    assert (out["soundEquilibrium"][good] <= out["soundFrozen"][good] * (1 + 2e-3)).all()
    rho = np.array([2.0, 5.0, 20.0]); T = np.array([300.0, 295.0, 310.0])
    fluid = cases.PureFluid("N2O")
    en = rho * fluid.energy(T, rho)
    s = np.zeros(3, dtype=abi.STATE_DTYPE); s["T"] = 280.0; s["p"] = 1e5
    out, ok, _ = oracle_flash(mb, rho[:, None], en, s)
    assert ok.all()
    assert out["T"] == pytest.approx(T, rel=1e-8)


def _two_cell(ul, ur, pl, pr, cl, cr, rhol=2.0, rhor=3.0, el=3e5, er=2e5):
    faces = np.zeros(3, dtype=abi.FACE_DTYPE)
    for i, (o, nb, kind, n) in enumerate([(0, 1, 0, (1, 0, 0)), (0, -1, 2, (-1, 0, 0)), (1, -1, 2, (1, 0, 0))]):
        faces[i]["owner"] = o; faces[i]["neighbour"] = nb; faces[i]["kind"] = kind; faces[i]["fixed"] = -1
        faces[i]["normal"] = n; faces[i]["area"] = 1e-4; faces[i]["distance"] = 1e-2
        faces[i]["ownerWeight"] = 0.5 if nb >= 0 else 1.0
    vol = np.array([1e-6, 1e-6])
    q = np.array([[rhol, rhol * ul, 0, 0, rhol * el], [rhor, rhor * ur, 0, 0, rhor * er]], float)
    st = np.zeros(2, dtype=abi.TSTATE_DTYPE)
    st["p"] = (pl, pr); st["T"] = 300; st["rho"] = (rhol, rhor); st["cv"] = 800; st["sound"] = (cl, cr)
    st["gasMass"] = (rhol, rhor)
    return faces, vol, q, st


def test_hll_single_face_by_hand_and_uniform_state():
    faces, vol, q, st = _two_cell(10.0, -5.0, 2e5, 1e5, 250.0, 300.0)
    ora = OracleTransport(config(2, 3), vol, faces)
    rhs, b = ora.rhs(q, st)
    # internal face by hand (HLL, Toro): S_L, S_R with waveFactor 1.1
    aL, aR = 1.1 * 250.0, 1.1 * 300.0
    SL, SR = min(0, 10 - aL, -5 - aR), max(0, 10 + aL, -5 + aR)
    FL = np.array([2 * 10, 2 * 10 * 10 + 2e5, 0, 0, 10 * (2 * 3e5 + 2e5)])
    FR = np.array([3 * -5, 3 * 25 + 1e5, 0, 0, -5 * (3 * 2e5 + 1e5)])
    F = (SR * FL - SL * FR + SL * SR * (q[1] - q[0])) / (SR - SL)
    # boundary faces: extrapolated, L = R = owner, normal -x on cell 0, +x on cell 1
    def hll_self(qc, u, pc, c, n):
        un = u * n; a = 1.1 * c
        sl, sr = min(0, un - a), max(0, un + a)
        f = qc * un
        f = f + np.array([0, pc * n, 0, 0, pc * un])
        return (sr * f - sl * f + sl * sr * 0) / (sr - sl)
    B0 = hll_self(q[0], 10.0, 2e5, 250.0, -1.0)
    B1 = hll_self(q[1], -5.0, 1e5, 300.0, 1.0)
    expect0 = -(F + B0) * 1e-4 / 1e-6
    expect1 = (F - B1) * 1e-4 / 1e-6
    assert rhs[0] == pytest.approx(expect0, rel=1e-9, abs=1e-6)
    assert rhs[1] == pytest.approx(expect1, rel=1e-9, abs=1e-6)
    assert b == pytest.approx((B0 + B1) * 1e-4, rel=1e-9, abs=1e-9)
    # a uniform state at rest has L(q) = 0 in the interior
    faces, vol, q, st = _two_cell(0.0, 0.0, 1e5, 1e5, 300.0, 300.0, 2.0, 2.0, 3e5, 3e5)
    rhs, b = OracleTransport(config(2, 3), vol, faces).rhs(q, st)
    assert np.abs(rhs[:, [0, 4]]).max() == 0.0


def test_mirror_symmetry_and_conservation():
    """Reflecting the x axis reverses x momentum and the x fluxes; sum rhs V + boundary = 0."""
    from helpers import random_graph
    faces, vol, q, st, fq, fst = random_graph(64, 3)
    cfg = config(len(q), len(faces), fixed=len(fq))
    ora = OracleTransport(cfg, vol, faces, fq, fst)
    rhs, b = ora.rhs(q, st)
    total = (rhs * vol[:, None]).sum(0) + b
    assert np.all(np.abs(total) <= 2e-11 * ((np.abs(rhs) * vol[:, None]).sum(0) + np.abs(b)))
    f2 = faces.copy(); f2["normal"][:, 0] *= -1
    q2 = q.copy(); q2[:, 1] *= -1
    fq2 = fq.copy(); fq2[:, 1] *= -1
    rhs2, b2 = OracleTransport(cfg, vol, f2, fq2, fst).rhs(q2, st)
    mirrored = rhs.copy(); mirrored[:, 1] *= -1
    assert np.allclose(rhs2, mirrored, rtol=1e-10, atol=1e-6 * np.abs(rhs).max())
