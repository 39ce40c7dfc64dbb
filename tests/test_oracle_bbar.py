"""Pins the oracle's B-bar element stiffness (hex8_ke_bbar) to the reference's
own Hex8SolidBbar outputs and to the properties fixed by its construction."""
import os

import numpy as np
import pytest

from oracle import h8_oracle as o

GOLD = os.path.join(os.path.dirname(__file__), "golden", "bbar_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("name", ["cube", "distorted", "brick_incomp"])
def test_bbar_ke_matches_reference(gold, name):
    E, nu = gold[f"ke/{name}/E_nu"]
    ke = o.hex8_ke_bbar(gold[f"ke/{name}/X"], o.d_solid3d(E, nu))
    ref = gold[f"ke/{name}/ke"]
    assert np.abs(ke - ref).max() <= 1e-13 * np.abs(ref).max()


def test_bbar_affine_energy_symmetry_rigid_modes():
    rng = np.random.default_rng(8)
    X = o.box_mesh(1.0, 1.0, 1.0, 1, 1, 1)[0][o.box_mesh(1.0, 1.0, 1.0, 1, 1, 1)[1][0]]
    X = X + rng.uniform(-0.1, 0.1, X.shape)
    D = o.d_solid3d(3.0, 0.45)
    ke = o.hex8_ke_bbar(X, D)
    assert np.allclose(ke, ke.T, atol=1e-13 * np.abs(ke).max())
    G = rng.normal(size=(3, 3))
    u = (X @ G.T).ravel()
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])
    pts, wts = o.gauss_points(2)
    V = sum(o.hex8_B(p, X)[1] * w for p, w in zip(pts, wts))
    assert np.isclose(u @ ke @ u, V * eps @ D @ eps, rtol=1e-12)
    W = np.array([[0, -1.0, 0.5], [1.0, 0, -0.3], [-0.5, 0.3, 0]])
    for r in (np.tile([1.0, 0, 0], 8), (X @ W.T).ravel()):
        assert np.abs(ke @ r).max() < 1e-12 * np.abs(ke).max()
    assert np.linalg.matrix_rank(ke, tol=1e-10 * np.abs(ke).max()) == 18
