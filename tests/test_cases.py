"""Synthetic case generator: mesh topology and state invariants."""
import numpy as np

from reactive_amd import cases


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
