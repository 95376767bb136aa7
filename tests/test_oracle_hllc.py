"""The oracle's HLLC face flux (the capillary branch's contact-preserving flux
with sigma = curvature = 0, reactiveBalancedCapillary.h:55-142) against the
reference's own faceFlux compiled into oracle/_ref, and against what the
mathematics fixes: a stationary contact carries no mass, a supersonic face is
the upwind flux, a strong pressure wave falls back to HLL, mirror symmetry."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import REF_PATH, oracle, p

V = C.c_void_p


def lib():
    L = oracle()
    L.oracle_hllc_face.argtypes = [V, V, V, V]
    return L


def ref():
    if not os.path.exists(REF_PATH):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    L = C.CDLL(REF_PATH)
    L.ref_hllc_face.argtypes = [V, V, V, V]
    return L


def face(L, fn, left, right, n):
    out = np.zeros(4)
    ok = getattr(L, fn)(p(np.asarray(left, float)), p(np.asarray(right, float)), p(np.asarray(n, float)), p(out))
    return ok, out


def random_states(rng, n):
    rho = np.exp(rng.uniform(np.log(0.5), np.log(1000), (n, 2)))
    u = rng.normal(0, 40, (n, 2, 3))
    u[rng.random(n) < 0.1] *= 20  # supersonic faces
    pr = np.exp(rng.uniform(np.log(1e4), np.log(5e7), (n, 2)))
    pr[rng.random(n) < 0.1, 1] *= 1e3  # strong pressure jumps (HLL fallback)
    a = rng.uniform(100, 1500, (n, 2))
    E = pr / 0.4 + 0.5 * rho * (u * u).sum(2)
    normal = rng.normal(size=(n, 3))
    normal /= np.linalg.norm(normal, axis=1)[:, None]
    left = np.concatenate([rho[:, :1], u[:, 0], pr[:, :1], a[:, :1], E[:, :1]], 1)
    right = np.concatenate([rho[:, 1:], u[:, 1], pr[:, 1:], a[:, 1:], E[:, 1:]], 1)
    return left, right, normal


def test_hllc_matches_reference_bitwise():
    L, R = lib(), ref()
    left, right, normal = random_states(np.random.default_rng(7), 4000)
    # axis-aligned normals too (exactly unit)
    normal[:500] = np.eye(3)[np.arange(500) % 3] * np.where(np.arange(500) % 2, 1, -1)[:, None]
    agree = 0
    for i in range(len(left)):
        o1, f1 = face(L, "oracle_hllc_face", left[i], right[i], normal[i])
        o2, f2 = face(R, "ref_hllc_face", left[i], right[i], normal[i])
        assert o1 == o2, i
        if o1:
            assert f1.tobytes() == f2.tobytes(), (i, f1, f2)
            agree += 1
    assert agree > 3000


def test_hllc_stationary_contact_carries_no_mass():
    """Equal pressure and zero velocity across a density jump: HLLC keeps the
    contact (zero mass flux); HLL would diffuse it."""
    L = lib()
    left = [900.0, 0, 0, 0, 4e6, 1200.0, 1e7]
    right = [50.0, 0, 0, 0, 4e6, 300.0, 1e7]
    ok, f = face(L, "oracle_hllc_face", left, right, [1.0, 0, 0])
    assert ok
    assert f[0] == 0.0 and f[1] == 0.0  # no advection of mass, momentum or energy
    assert f[2] == 4e6 and f[3] == 0.0   # pressure momentum p, no pressure work


def test_hllc_supersonic_is_upwind():
    L = lib()
    left = [2.0, 3000.0, 5, -7, 1e5, 300.0, 2.5e5 + 0.5 * 2 * (3000 ** 2 + 25 + 49)]
    right = [1.0, 2900.0, 0, 0, 2e5, 350.0, 5e5 + 0.5 * 2900 ** 2]
    ok, f = face(L, "oracle_hllc_face", left, right, [1.0, 0, 0])
    assert ok
    # mass flux rho u_n from the left state alone
    assert f[0] * left[0] == pytest.approx(left[0] * left[1], rel=1e-15) and f[1] == 0.0
    assert f[2] == left[4] and f[3] == left[4] * left[1]


def test_hllc_mirror_symmetry():
    """Swapping sides and flipping the normal and the normal velocity reverses the flux."""
    L = lib()
    rng = np.random.default_rng(3)
    left, right, normal = random_states(rng, 300)
    for i in range(len(left)):
        ok1, f1 = face(L, "oracle_hllc_face", left[i], right[i], normal[i])
        ok2, f2 = face(L, "oracle_hllc_face", right[i], left[i], -normal[i])
        assert ok1 == ok2
        if ok1:
            # mass flux through the face: advectL rho_L + advectR rho_R, antisymmetric
            m1 = f1[0] * left[i][0] + f1[1] * right[i][0]
            m2 = f2[0] * right[i][0] + f2[1] * left[i][0]
            assert m1 == pytest.approx(-m2, rel=1e-9, abs=1e-9 * (abs(m1) + 1))
            assert f1[2] == pytest.approx(f2[2], rel=1e-9)  # pressure momentum (multiplies n)
