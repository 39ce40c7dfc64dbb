"""h8pcg: MI355X (gfx950) matrix-free H8 linear-elastic solver.

A drop-in for fea_engine's Hex8 assemble / Jacobi-PCG solve path: the host API
mirrors fea_engine (FESystem, box_mesh, Hex8Solid3D, Material, D_solid3d,
coefficients.by_element), and the element, operator, solver, recovery and SIMP
arithmetic runs in libh8pcg.so (hand-written HIP kernels) on the GPU; the host
builds meshes, loads, the constraint map and the rigid-mode basis.
"""
from .material import Material, D_solid3d
from .mesh import Mesh, box_mesh, structured_grid
from .elements import Hex8Solid3D, Hex8SolidBbar
from .hex20 import Hex20Solid3D, box_mesh_hex20
from . import coefficients
from .fields import FEField
from .facet_loads import list_facets, boundary_faces, face_load_shares
from .solver import FESystem, fesystem_solve_static_torch
from ._lib import NotSPDError, H8LibraryMissing
from .batched import solve_static_batched, solve_many

__all__ = ["Material", "D_solid3d", "Mesh", "box_mesh", "structured_grid", "Hex8Solid3D", "Hex8SolidBbar", "Hex20Solid3D", "box_mesh_hex20", "coefficients",
           "FEField", "list_facets", "boundary_faces", "face_load_shares", "FESystem", "fesystem_solve_static_torch",
           "NotSPDError",
           "H8LibraryMissing", "solve_static_batched", "solve_many"]
__version__ = "0.1.0"
