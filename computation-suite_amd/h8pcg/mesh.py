"""Hex8 box meshes and node selection (host side).

Mirrors the subset of fe/mesh.py the H8 path uses: ``Mesh`` with ``nodes``,
``elements``, ``dim``, node sets, ``select_nodes`` / ``nodes_on_plane``
(fe/mesh.py:196-332) and ``box_mesh`` (fe/mesh.py:716-733,1270-1289).  Meshes
are generated vectorised; a box mesh records its structure so the solver can
use the implicit-connectivity device operator, and ``structured_grid`` recovers
that structure from any Mesh whose arrays match a box mesh.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    nodes: np.ndarray      # (n_nodes, dim)
    elements: np.ndarray   # (n_elements, 8)
    dim: int
    node_sets: dict = field(default_factory=dict)
    structure: tuple | None = None   # (nx, ny, nz, Lx, Ly, Lz) for box meshes

    # ---- node sets / selection (the contract of fe/mesh.py:206-264) -------------
    # A node set is a sorted array of unique node ids stored under a name; the
    # solver's node arguments accept those names.  Errors are ValueError.
    def add_node_set(self, name, node_ids):
        if not (isinstance(name, str) and name):
            raise ValueError("add_node_set: a node set needs a non-empty string name.")
        ids = np.asarray(node_ids).ravel()
        if ids.size == 0:
            raise ValueError(f"add_node_set({name!r}): the set would be empty (a selection that matched no "
                             f"node; check the coordinate and tol).")
        if ids.dtype.kind == "f" and np.array_equal(ids, np.floor(ids)):
            ids = ids.astype(np.int64)
        elif ids.dtype.kind not in "iu":
            raise ValueError(f"add_node_set({name!r}): node ids have to be whole numbers.")
        ids = np.unique(ids.astype(np.int64))
        if ids[0] < 0 or ids[-1] >= len(self.nodes):
            raise ValueError(f"add_node_set({name!r}): ids run outside 0..{len(self.nodes) - 1}.")
        self.node_sets[name] = ids
        return self

    def node_set(self, name):
        if name not in self.node_sets:
            raise ValueError(f"unknown node set {name!r} (defined: {sorted(self.node_sets)})")
        return self.node_sets[name]

    def select_nodes(self, x=None, y=None, z=None, tol=1e-9, name=None):
        """Ids of the nodes matching every given coordinate: a value (within tol)
        or a closed interval (lo, hi) widened by tol; registered as a node set
        when `name` is given."""
        coords = self.nodes
        keep = np.ones(len(coords), dtype=bool)
        for axis, label, want in ((0, "x", x), (1, "y", y), (2, "z", z)):
            if want is None:
                continue
            if axis >= coords.shape[1]:
                raise ValueError(f"select_nodes: {label}= given for a mesh with {coords.shape[1]} "
                                 f"coordinate(s).")
            c = coords[:, axis]
            if np.ndim(want) == 0:
                keep &= np.abs(c - want) < tol
                continue
            lo, hi = want
            if hi < lo:
                raise ValueError(f"select_nodes: the interval {label}=({lo}, {hi}) is reversed.")
            keep &= (c >= lo - tol) & (c <= hi + tol)
        ids = np.flatnonzero(keep)
        if name is not None:
            self.add_node_set(name, ids)
        return ids

    def nodes_on_plane(self, axis, value, tol=1e-9):
        return np.where(np.abs(self.nodes[:, axis] - value) < tol)[0]

    @property
    def n_nodes(self):
        return len(self.nodes)

    @property
    def n_elements(self):
        return len(self.elements)


def _box_arrays(Lx, Ly, Lz, nx, ny, nz):
    dx, dy = Lx / nx, Ly / ny
    n2 = (nx + 1) * (ny + 1)
    i = np.repeat(np.arange(nx + 1), ny + 1)
    j = np.tile(np.arange(ny + 1), nx + 1)
    z = np.linspace(0.0, Lz, nz + 1)
    nodes = np.empty((n2 * (nz + 1), 3))
    nodes[:, 0] = np.tile(i * dx, nz + 1)
    nodes[:, 1] = np.tile(j * dy, nz + 1)
    nodes[:, 2] = np.repeat(z, n2)
    ie = np.repeat(np.arange(nx), ny)
    je = np.tile(np.arange(ny), nx)
    base = ie * (ny + 1) + je
    quad = np.stack([base, base + (ny + 1), base + (ny + 1) + 1, base + 1], axis=1)
    k = np.repeat(np.arange(nz), nx * ny)[:, None] * n2
    quad = np.tile(quad, (nz, 1))
    elements = np.concatenate([quad + k, quad + k + n2], axis=1).astype(np.int64)
    return nodes, elements


def box_mesh(Lx, Ly, Lz, nx, ny, nz):
    """Structured Hex8 box [0,Lx]x[0,Ly]x[0,Lz] with nx x ny x nz elements:
    node(i,j,k) = k(nx+1)(ny+1) + i(ny+1) + j, element e(i,j,k) = k*nx*ny + i*ny + j
    (fe/mesh.py:716-733, 1270-1289)."""
    for n in (nx, ny, nz):
        if int(n) != n or n < 1:
            raise ValueError("box_mesh: nx, ny, nz must be positive integers")
    nodes, elements = _box_arrays(float(Lx), float(Ly), float(Lz), int(nx), int(ny), int(nz))
    return Mesh(nodes, elements, dim=3, structure=(int(nx), int(ny), int(nz), float(Lx), float(Ly), float(Lz)))


def structured_grid(mesh, rtol=1e-12):
    """(nx, ny, nz, h) if `mesh` is a box mesh whose elements are all translates of
    one brick (so one K0 serves every element), else None.  h = element corner
    offsets (8,3) of element 0."""
    nodes = np.asarray(mesh.nodes, dtype=float)
    elems = np.asarray(mesh.elements)
    if nodes.ndim != 2 or nodes.shape[1] != 3 or elems.ndim != 2 or elems.shape[1] != 8:
        return None
    st = getattr(mesh, "structure", None)
    if st is not None:
        nx, ny, nz = st[:3]
    else:
        nx = len(np.unique(np.round(nodes[:, 0], 12))) - 1
        ny = len(np.unique(np.round(nodes[:, 1], 12))) - 1
        nz = len(np.unique(np.round(nodes[:, 2], 12))) - 1
        if min(nx, ny, nz) < 1:
            return None
    if len(nodes) != (nx + 1) * (ny + 1) * (nz + 1) or len(elems) != nx * ny * nz:
        return None
    _, ref_e = _box_arrays(1.0, 1.0, 1.0, nx, ny, nz)
    if not np.array_equal(elems, ref_e):
        return None
    rel = nodes[elems] - nodes[elems[:, :1]]
    h = rel[0]
    scale = max(np.abs(h).max(), 1e-300)
    if np.abs(rel - h[None]).max() > rtol * scale * 16:
        return None
    # the brick must be axis-aligned with positive extents for the shared-K0 operator
    return nx, ny, nz, nodes[elems[0]]
