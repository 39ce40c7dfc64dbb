"""Pins the oracle's stress recovery (oracle.h8_oracle.recover / von_mises):
the reference's own outputs on the committed fixture, closed forms for uniform
strain and rigid motion, and von Mises of uniaxial stress and pure shear."""
import os

import numpy as np
import pytest

from oracle import h8_oracle as o

GOLD = os.path.join(os.path.dirname(__file__), "golden", "recovery_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _case(gold, name):
    nx, ny, nz = (int(v) for v in gold[name + "/dims"])
    Lx, Ly, Lz = (float(v) for v in gold[name + "/lengths"])
    nodes, elems = o.box_mesh(Lx, Ly, Lz, nx, ny, nz)
    E = gold[name + "/E"]
    scale = np.broadcast_to(E, (len(elems),)) if E.ndim == 0 else E
    return nodes, elems, scale, gold[name + "/U"]


@pytest.mark.parametrize("name", ["het", "const"])
def test_recover_matches_reference_fixture(gold, name):
    nodes, elems, scale, U = _case(gold, name)
    D1 = o.d_solid3d(1.0, 0.3)
    sn, se = o.recover(nodes, elems, U, D1, scale)
    en, ee = o.recover(nodes, elems, U, None)
    for got, key in ((sn.ravel(), "stress_nodes"), (se, "stress_elements"), (en.ravel(), "strain_nodes"),
                     (ee, "strain_elements"), (o.von_mises(sn), "vm_nodes"), (o.von_mises(se), "vm_elements")):
        ref = gold[f"{name}/{key}"]
        assert got.shape == ref.shape, key
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), key


def _affine_field(nodes, G, c):
    """u(x) = G x + c at every node, flattened (3 per node)."""
    return (nodes @ G.T + c).ravel()


def test_uniform_strain_gives_D_eps_everywhere():
    rng = np.random.default_rng(3)
    nodes, elems = o.box_mesh(3.0, 1.0, 2.0, 3, 2, 4)
    G = rng.normal(size=(3, 3))
    U = _affine_field(nodes, G, rng.normal(size=3))
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])
    D = o.d_solid3d(7.0, 0.27)
    sn, se = o.recover(nodes, elems, U, D)
    sig = D @ eps
    assert np.allclose(sn, sig[None, :], rtol=0, atol=1e-12 * np.abs(sig).max())
    assert np.allclose(se, sig[None, :], rtol=0, atol=1e-12 * np.abs(sig).max())
    en, _ = o.recover(nodes, elems, U, None)
    assert np.allclose(en, eps[None, :], rtol=0, atol=1e-12 * np.abs(eps).max())


def test_rigid_motion_gives_zero_stress():
    rng = np.random.default_rng(4)
    nodes, elems = o.box_mesh(2.0, 2.0, 1.0, 2, 3, 2)
    w = rng.normal(size=3)
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])   # infinitesimal rotation
    U = _affine_field(nodes, W, rng.normal(size=3))
    sn, se = o.recover(nodes, elems, U, o.d_solid3d(1.0, 0.3))
    assert np.abs(sn).max() < 1e-13 and np.abs(se).max() < 1e-13


def test_von_mises_closed_forms():
    assert np.isclose(o.von_mises(np.array([5.0, 0, 0, 0, 0, 0])), 5.0)
    assert np.isclose(o.von_mises(np.array([0, -3.0, 0, 0, 0, 0])), 3.0)
    assert np.isclose(o.von_mises(np.array([0, 0, 0, 0, 2.0, 0])), np.sqrt(3) * 2.0)
    assert np.isclose(o.von_mises(np.array([4.0, 4.0, 4.0, 0, 0, 0])), 0.0)   # hydrostatic


def test_projection_of_trilinear_field_by_hand():
    """One element, u_x = x*y*z (trilinear, exact in H8): eps_xx(p) = y_p z_p at the
    Gauss points, so the corner value is sum_p N_a(p) y_p z_p (weights 1, sum_p N_a = 1)."""
    nodes, elems = o.box_mesh(2.0, 2.0, 2.0, 1, 1, 1)
    X = nodes - 1.0                                     # centred: natural = physical coordinates
    U = np.zeros(3 * len(nodes))
    U[0::3] = X[:, 0] * X[:, 1] * X[:, 2]
    _, en = o.recover(nodes, elems, U, None)
    nn, _ = o.recover(nodes, elems, U, None)
    pts, _ = o.gauss_points(2)
    for a in range(8):
        node = elems[0][a]
        want = sum(o.hex8_shape_and_derivs(*p)[0][a] * p[1] * p[2] for p in pts)
        assert np.isclose(nn[node, 0], want, atol=1e-14)
    assert abs(en[0, 0]) < 1e-14                        # volume average of y z is 0
