"""Pins the oracle's element arithmetic (hex8_shape_and_derivs, hex8_B, hex8_ke,
d_solid3d) with checks that a dropped term, a wrong sign, a wrong Voigt row or a
transposed operand would fail."""
import numpy as np
import pytest

from oracle import h8_oracle as o

UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)


def distorted(seed=0, amp=0.15):
    rng = np.random.default_rng(seed)
    return UNIT * np.array([1.3, 0.8, 1.1]) + amp * rng.uniform(-1, 1, (8, 3))


def rigid_modes(X):
    """6 rigid-body displacement fields (3 translations, 3 infinitesimal rotations)."""
    modes = []
    for c in range(3):
        v = np.zeros((8, 3)); v[:, c] = 1.0; modes.append(v.ravel())
    for axis in np.eye(3):
        modes.append(np.cross(axis, X).ravel())
    return np.array(modes)


def test_shape_functions_partition_and_kronecker():
    rng = np.random.default_rng(1)
    for p in rng.uniform(-1, 1, (5, 3)):
        N, dN = o.hex8_shape_and_derivs(*p)
        assert abs(N.sum() - 1) < 1e-14
        assert np.abs(dN.sum(axis=1)).max() < 1e-14
        # derivative vs central difference
        h = 1e-6
        for ax in range(3):
            e = np.zeros(3); e[ax] = h
            fd = (o.hex8_shape_and_derivs(*(p + e))[0] - o.hex8_shape_and_derivs(*(p - e))[0]) / (2 * h)
            assert np.abs(fd - dN[ax]).max() < 1e-9
    corners = np.stack([o.XI_I, o.ETA_I, o.ZETA_I], axis=1)
    for a, c in enumerate(corners):
        N, _ = o.hex8_shape_and_derivs(*c)
        assert np.allclose(N, np.eye(8)[a], atol=1e-15)


@pytest.mark.parametrize("X", [UNIT, distorted(0), distorted(3)])
def test_ke_symmetric_rank18_rigid_null(X):
    D = o.d_solid3d(2.1e5, 0.3)
    ke = o.hex8_ke(X, D)
    assert np.abs(ke - ke.T).max() < 1e-10 * np.abs(ke).max()
    R = rigid_modes(X)
    assert np.abs(ke @ R.T).max() < 1e-9 * np.abs(ke).max()
    w = np.linalg.eigvalsh(ke)
    assert np.sum(w > 1e-8 * w.max()) == 18


def test_ke_two_point_rule_is_exact_on_brick():
    # the trilinear brick's integrand is quadratic per axis: 2 points are exact
    X = UNIT * np.array([2.0, 0.5, 1.5])
    D = o.d_solid3d(1.0, 0.25)
    assert np.abs(o.hex8_ke(X, D, 2) - o.hex8_ke(X, D, 4)).max() < 1e-13


def test_ke_linear_in_E():
    X = distorted(5)
    k1 = o.hex8_ke(X, o.d_solid3d(1.0, 0.3))
    k7 = o.hex8_ke(X, o.d_solid3d(7.0, 0.3))
    assert np.abs(k7 - 7 * k1).max() < 1e-12 * np.abs(k7).max()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_uniform_strain_energy(seed):
    """u = G x gives eps = sym(G) everywhere: u^T ke u = V * eps^T D eps, and
    the lame form lambda*tr^2 + 2*mu*eps:eps equals eps^T D eps (pins D too)."""
    rng = np.random.default_rng(seed)
    X = distorted(seed, 0.1)
    E, nu = 3.0, 0.27
    D = o.d_solid3d(E, nu)
    G = rng.normal(size=(3, 3))
    u = (X @ G.T).ravel()
    eps = 0.5 * (G + G.T)
    ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])
    pts, wts = o.gauss_points(3)
    V = sum(o.hex8_B(p, X)[1] * w for p, w in zip(pts, wts))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    e_lame = lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps)
    assert abs(ev @ D @ ev - e_lame) < 1e-12 * abs(e_lame)
    ke = o.hex8_ke(X, D)
    assert abs(u @ ke @ u - V * e_lame) < 1e-10 * abs(V * e_lame)


def test_patch_test_distorted_mesh():
    """A linear displacement prescribed on the boundary of a distorted 2x2x2 mesh
    is reproduced exactly at the interior node (fe/tests/test_benchmarks_convergence.py:144-174)."""
    nodes, elems = o.box_mesh(1.0, 1.0, 1.0, 2, 2, 2)
    nodes = nodes.copy()
    centre = np.where(np.all(np.abs(nodes - 0.5) < 1e-9, axis=1))[0][0]
    nodes[centre] += np.array([0.07, -0.05, 0.04])
    D = o.d_solid3d(2.1e5, 0.3)
    K = o.assemble_csr(nodes, elems, D)
    G = np.array([[1e-3, 2e-4, 1e-4], [3e-4, -5e-4, 2e-4], [1e-4, 4e-4, 8e-4]])
    exact = nodes @ G.T + np.array([1e-3, -2e-3, 5e-4])
    bnd = [n for n in range(len(nodes)) if n != centre]
    fixed = [3 * n + c for n in bnd for c in range(3)]
    vals = {3 * n + c: exact[n, c] for n in bnd for c in range(3)}
    U = o.solve_direct(K, np.zeros(K.shape[0]), fixed, vals)
    assert np.abs(U.reshape(-1, 3) - exact).max() < 1e-10 * np.abs(exact).max()
