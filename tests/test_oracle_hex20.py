"""Pins of the oracle's Hex20 restatement (fe/elements/solids.py:331-406), by
the checks the reference's own tests make of the element
(fe/tests/test_higher_order_elements.py:175-230: partition of unity and the
isoparametric identity, the constant-strain patch test, six rigid-body modes)
and by mathematics: Kronecker property at the nodes, derivatives equal to
central differences of N, exact integration of the brick by the 3x3x3 rule,
symmetry and rank 54, and the library's box numbering against a box built as
the reference's tests build quadratic meshes."""
import numpy as np

from oracle import h8_oracle as o

NAT = np.stack([o.XI20, o.ETA20, o.ZETA20], axis=1)


def _distorted_hex20(seed=0):
    """A distorted Hex20 with straight edges (mid-edge nodes at the midpoints),
    so linear fields are represented exactly."""
    rng = np.random.default_rng(seed)
    corners = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                       dtype=float) * np.array([1.3, 0.9, 1.1])
    corners += 0.15 * (rng.random((8, 3)) - 0.5)
    X = np.zeros((20, 3))
    X[:8] = corners
    for a in range(8, 20):
        N8, _ = o.hex8_shape_and_derivs(*NAT[a])
        X[a] = N8 @ corners
    return X


def test_partition_of_unity_kronecker_and_isoparametric_identity():
    rng = np.random.default_rng(1)
    X = _distorted_hex20()
    for p in rng.uniform(-1, 1, (10, 3)):
        N, dN = o.hex20_shape_and_derivs(*p)
        assert abs(N.sum() - 1.0) < 1e-14 and np.abs(dN.sum(axis=1)).max() < 1e-13
        # isoparametric identity: sum N_a x_a is the point the trilinear map sends p to
        N8, _ = o.hex8_shape_and_derivs(*p)
        assert np.allclose(N @ X, N8 @ X[:8], atol=1e-13)
    for b in range(20):
        N, _ = o.hex20_shape_and_derivs(*NAT[b])
        assert np.allclose(N, np.eye(20)[b], atol=1e-14)


def test_derivatives_are_central_differences_of_N():
    rng = np.random.default_rng(2)
    h = 1e-6
    for p in rng.uniform(-1, 1, (5, 3)):
        _, dN = o.hex20_shape_and_derivs(*p)
        for d in range(3):
            e = np.eye(3)[d] * h
            fd = (o.hex20_shape_and_derivs(*(p + e))[0] - o.hex20_shape_and_derivs(*(p - e))[0]) / (2 * h)
            assert np.abs(fd - dN[d]).max() < 1e-8


def test_stiffness_symmetry_rigid_modes_rank_and_patch_test():
    D = o.d_solid3d(210e9, 0.3)
    X = _distorted_hex20(3)
    ke = o.hex20_ke(X, D)
    assert np.abs(ke - ke.T).max() <= 1e-12 * np.abs(ke).max()
    # six rigid-body modes in the null space, rank 60 - 6
    R = np.zeros((60, 6))
    for c in range(3):
        R[c::3, c] = 1.0
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    R[0::3, 3], R[1::3, 3] = -y, x
    R[1::3, 4], R[2::3, 4] = -z, y
    R[0::3, 5], R[2::3, 5] = z, -x
    assert np.abs(ke @ R).max() <= 1e-9 * np.abs(ke).max()
    assert np.linalg.matrix_rank(ke, tol=1e-8 * np.abs(ke).max()) == 54
    # constant strain: u = A x gives energy V eps^T D eps (the patch test of one element)
    A = np.random.default_rng(4).standard_normal((3, 3)) * 1e-3
    u = (X @ A.T).ravel()
    eps = np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[1, 2] + A[2, 1], A[0, 2] + A[2, 0]])
    pts, wts = o.gauss_points(3)
    V = sum(w * o.hex20_B(*p, X)[1] for p, w in zip(pts, wts))
    assert abs(u @ ke @ u - V * eps @ D @ eps) <= 1e-10 * abs(V * eps @ D @ eps)
    # the strain at every Gauss point is the constant one
    for p in pts[:5]:
        B, _ = o.hex20_B(*p, X)
        assert np.allclose(B @ u, eps, atol=1e-14)


def test_brick_is_integrated_exactly_by_the_3x3x3_rule():
    D = o.d_solid3d(1.0, 0.3)
    X = np.zeros((20, 3))
    X[:] = (NAT + 1.0) * 0.5 * np.array([2.0, 1.0, 0.5])
    assert np.allclose(o.hex20_ke(X, D, order=3), o.hex20_ke(X, D, order=4), rtol=0, atol=1e-13)


def test_library_box_numbering_matches_the_midpoint_construction():
    """The library's box_mesh_hex20 (host mirror) and the reference-style
    construction give the same element node coordinates, element by element."""
    from h8pcg.hex20 import _hex20_arrays
    nx, ny, nz = 3, 2, 2
    nodes_lib, elems_lib = _hex20_arrays(3.0, 2.0, 1.5, nx, ny, nz)
    nodes_ref, elems_ref = o.hex20_box_by_midpoints(3.0, 2.0, 1.5, nx, ny, nz)
    assert len(nodes_lib) == len(nodes_ref)
    assert np.allclose(nodes_lib[elems_lib], nodes_ref[elems_ref], atol=1e-12)
    # every node used, none twice at one place
    assert len(np.unique(elems_lib)) == len(nodes_lib)
    assert len(np.unique(np.round(nodes_lib, 9), axis=0)) == len(nodes_lib)
