"""Builders for the benchmark workloads (box-mesh cantilevers).

cantilever(nx, ny, nz, ...) -> assembled FESystem: unit-cube H8 elements,
x = 0 face clamped, a uniform -z traction of `total` on the x = nx face as a
consistent facet load (or an even nodal split), per-element Young's modulus
E_elem (or one E).  batched_cantilevers(seed, n) -> the uneven multi-instance
batch: nx drawn from {8, 12, ..., 48}, ny = nz = nx/4, 4..16 load cases each
(unit-total tip traction along a normalised Gaussian direction), per-element
E = 10**U(9, 11).
"""
from __future__ import annotations

import numpy as np

from .elements import Hex8Solid3D
from .facet_loads import boundary_faces
from .material import D_solid3d, Material
from .mesh import box_mesh
from .solver import FESystem


def cantilever(nx, ny, nz, E=210e9, nu=0.3, E_elem=None, load="facet", total=-1e6, device="cuda"):
    mesh = box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = FESystem(mesh, Hex8Solid3D(), backend="hip", device=device)
    if E_elem is None:
        fs.assemble_stiffness(D_solid3d(Material(E=E, nu=nu)))
    else:
        fs.assemble_stiffness(D_solid3d(Material(E=1.0, nu=nu)), element_scale=np.asarray(E_elem, dtype=float))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    if load == "facet":
        fs.add_consistent_facet_load(Hex8Solid3D(), boundary_faces(mesh, 0, float(nx)), 2, total / (ny * nz))
    elif load == "nodal":
        fs.add_nodal_force(mesh.nodes_on_plane(0, float(nx)), 2, total)
    return fs


def tip_traction_loads(fs, nx, ny, nz, k, rng):
    """(n_dof, k) load cases: unit-total traction on the x = nx face along g/|g|, g ~ N(0, I3)."""
    faces = boundary_faces(fs.mesh, 0, float(nx))
    ids = np.asarray(faces, dtype=int)
    from .facet_loads import face_load_shares_batched
    shares = face_load_shares_batched(np.asarray(fs.mesh.nodes)[ids], 1.0 / (ny * nz))
    F = np.zeros((fs.n_dof, k))
    for c in range(k):
        g = rng.normal(size=3)
        g /= np.linalg.norm(g)
        for d in range(3):
            np.add.at(F[:, c], 3 * ids.ravel() + d, g[d] * shares.ravel())
    return F


def batched_cantilevers(seed=4, n=256, device="cuda"):
    """[(FESystem, F (n_dof, k))] for the uneven multi-instance batch."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = 4 * int(rng.integers(2, 13))
        ny = nz = nx // 4
        k = int(rng.integers(4, 17))
        E = 10 ** rng.uniform(9, 11, nx * ny * nz)
        fs = cantilever(nx, ny, nz, E_elem=E, load="none", device=device)
        out.append((fs, tip_traction_loads(fs, nx, ny, nz, k, rng)))
    return out
