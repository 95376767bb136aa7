"""Synthetic case generator: mesh topology and state invariants."""
import ctypes as C

import numpy as np

from helpers import oracle, p
from reactive_amd import cases, model


def test_cartesian_mesh_counts_and_orientation():
    for n in (3, 5):
        faces, volume = cases.cartesian_faces(n, 1e-3)
        assert len(faces) == 3 * n * n * (n + 1) and len(volume) == n ** 3
        internal = faces[faces["neighbour"] >= 0]
        assert len(internal) == 3 * n * n * (n - 1)
        assert (internal["owner"] < internal["neighbour"]).all()
        # internal faces are ordered by owner then neighbour (upper-triangular)
        key = internal["owner"] * n ** 3 + internal["neighbour"]
        assert (np.diff(key) > 0).all()
        # each cell has exactly six faces; normals are unit axis vectors
        count = np.bincount(faces["owner"], minlength=n ** 3) + np.bincount(internal["neighbour"], minlength=n ** 3)
        assert (count == 6).all()
        assert np.allclose(np.linalg.norm(faces["normal"], axis=1), 1.0)


def test_equilibrium_states_are_consistent():
    faces, volume, q, seeds, info = cases.pure_n2o_cube(6, seed=4)
    assert 0.2 < info["two_phase_fraction"] < 0.8
    e = cases.internal_energy(q, 1)
    assert np.allclose(e, q[:, 0] * seeds["e"], rtol=1e-12)
    assert (seeds["liquidMass"][:, 0] <= q[:, 0]).all()


def test_mixture_energy_matches_the_oracle_pr_state():
    """The numpy PR mixture energy that seeds binary cases equals the oracle's
    single-phase PR state at (T, rho, Y) (oracle_pr_trho, reactiveDevicePR.h:840-930)."""
    mb = model.model_bytes(model.build_model(gas_species=("N2O", "N2")))
    img = C.create_string_buffer(mb, len(mb))
    mix = cases.MixtureFluid(("N2O", "N2"))
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(200):
        T, rho, y = rng.uniform(200, 320), np.exp(rng.uniform(np.log(0.5), np.log(60))), rng.uniform(0, .3)
        Y = np.array([1 - y, y])
        out = np.zeros(40)
        if oracle().oracle_pr_trho(img, 0, T, rho, p(Y), p(out)) == 0:
            checked += 1
            assert abs(mix.energy(np.array([T]), rho, Y)[0] - out[5]) <= 4e-15 * abs(out[5])
            assert abs(mix.pressure(np.array([T]), rho, Y)[0] - out[1]) <= 1e-13 * abs(out[1])
    assert checked > 150


def test_binary_cube_composition():
    faces, volume, q, seeds, info = cases.binary_cube(5, seed=2)
    rho = q[:, :2].sum(1)
    assert (q[:, 1] / rho <= 0.3).all() and (q[:, 1] >= 0).all()
    e = cases.internal_energy(q, 2)
    assert np.allclose(e, rho * seeds["e"], rtol=1e-12)
