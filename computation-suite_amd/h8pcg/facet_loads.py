"""Consistent facet loads on Hex8 faces (host side, tiny).

FACE_TABLE is the Hex8 face-to-local-node table of fe/facet_loads.py:172-179;
face_load_shares integrates N * t over a bilinear face with a 2x2 Gauss rule and
the surface measure sqrt(det(J J^T)) (fe/facet_loads.py:283-325)."""
from __future__ import annotations

import numpy as np

# (local nodes, family) -- faces: z-, z+, y-, x+, y+, x-
FACE_TABLE = [((0, 3, 2, 1), "quad4"), ((4, 5, 6, 7), "quad4"), ((0, 1, 5, 4), "quad4"),
              ((1, 2, 6, 5), "quad4"), ((2, 3, 7, 6), "quad4"), ((3, 0, 4, 7), "quad4")]

_XA = np.array([-1.0, 1.0, 1.0, -1.0])
_EA = np.array([-1.0, -1.0, 1.0, 1.0])


# natural coordinates of the Hex8 corners (fe/elements/solids.py:135-137 order)
_HEX8_NAT = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def _face_map(loc):
    """(a, b) -> (xi, eta, zeta) of a face: the bilinear interpolation of its
    corners' natural coordinates, listed in `loc` order, which is affine on a face
    of the parent cube.  Facet corner k sits at (_XA[k], _EA[k])."""
    C = _HEX8_NAT[list(loc)]

    def nat(a, b):
        N = 0.25 * (1 + a * _XA) * (1 + b * _EA)
        return tuple(float(v) for v in N @ C)
    return nat


def quad4_shape(p):
    """N (4,) and dN/d(a, b) (2, 4) of the bilinear facet at (a, b)."""
    a, b = p
    N = 0.25 * (1 + a * _XA) * (1 + b * _EA)
    return N, np.vstack([0.25 * _XA * (1 + b * _EA), 0.25 * _EA * (1 + a * _XA)])


_FACET_SHAPE_FNS = {"quad4": quad4_shape}
_FACETS = [(loc, fam, _face_map(loc)) for loc, fam in FACE_TABLE]


def list_facets(formulation):
    """(local node indices, family, natural-coordinate map) of every face of a
    Hex8 parent (fe/facet_loads.py:191-211): `conn[list(idx)]` gives the global
    node tuple add_consistent_facet_load expects.  ValueError for any other
    element type, as the reference (whose table is keyed by the exact class)."""
    from .elements import Hex8Solid3D
    if type(formulation) is not Hex8Solid3D:
        raise ValueError(f"facet_loads: no facet table for {type(formulation).__name__} -- this backend's "
                         f"facet table covers Hex8Solid3D.")
    return list(_FACETS)


def _gauss(n):
    p, w = np.polynomial.legendre.leggauss(n)
    return [(a, b) for a in p for b in p], [wa * wb for wa in w for wb in w]


def face_load_shares(face_coords, traction, quad_order=2):
    """Shares (4,) of a uniform scalar traction on one bilinear face."""
    X = np.asarray(face_coords, dtype=float)
    if X.shape != (4, 3):
        raise ValueError(f"a Hex8 face has 4 nodes with 3 coordinates, got shape {X.shape}")
    n = max(1, int(np.ceil((quad_order + 1) / 2)))
    pts, wts = _gauss(n)
    share = np.zeros(4)
    for (a, b), w in zip(pts, wts):
        N = 0.25 * (1 + a * _XA) * (1 + b * _EA)
        dN = np.vstack([0.25 * _XA * (1 + b * _EA), 0.25 * _EA * (1 + a * _XA)])
        J = dN @ X
        share += N * traction * np.sqrt(np.linalg.det(J @ J.T)) * w
    return share


def face_load_shares_batched(face_coords, traction, quad_order=2):
    """Vectorised face_load_shares for (k,4,3) faces."""
    X = np.asarray(face_coords, dtype=float)
    n = max(1, int(np.ceil((quad_order + 1) / 2)))
    pts, wts = _gauss(n)
    share = np.zeros(X.shape[:2])
    for (a, b), w in zip(pts, wts):
        N = 0.25 * (1 + a * _XA) * (1 + b * _EA)
        dN = np.vstack([0.25 * _XA * (1 + b * _EA), 0.25 * _EA * (1 + a * _XA)])
        J = np.einsum('rn,knc->krc', dN, X)
        meas = np.sqrt(np.linalg.det(np.einsum('krc,ksc->krs', J, J)))
        share += N[None, :] * (traction * meas * w)[:, None]
    return share


def boundary_faces(mesh, axis, value, tol=1e-9):
    """Global node tuples of every Hex8 face whose 4 nodes lie on the plane
    coordinate[axis] == value, in FACE_TABLE order (the facets argument of
    FESystem.add_consistent_facet_load)."""
    elems = np.asarray(mesh.elements)
    on = np.abs(np.asarray(mesh.nodes)[:, axis] - value) < tol
    out = []
    for loc, _ in FACE_TABLE:
        idx = elems[:, list(loc)]
        hit = np.all(on[idx], axis=1)
        out.extend(tuple(int(v) for v in row) for row in idx[hit])
    return out
