"""Hex8Solid3D: the 8-node trilinear solid element (fe/elements/solids.py:130-158),
and Hex8SolidBbar, its B-bar (mean-dilatation) form (fe/elements/solids.py:225-272).

The element descriptor mirrors the reference class (n_nodes, dofs_per_node,
dim, gauss_order, DOF names).  ``stiffness`` / ``stiffness_batched`` evaluate
sum_gp B^T D B detJ w (2x2x2 Gauss rule, signed detJ, fe/elements/base.py:
494-514) on the GPU through h8_ke_batched_f64."""
from __future__ import annotations

import numpy as np

_XI = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=float)
_ETA = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=float)
_ZETA = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=float)


class Hex8Solid3D:
    n_nodes, dofs_per_node, dim, gauss_order = 8, 3, 3, 2

    @staticmethod
    def local_dof_names():
        return ("ux", "uy", "uz")

    def dof_index(self, name):
        names = self.local_dof_names()
        aliases = {"x": 0, "y": 1, "z": 2, "u": 0, "v": 1, "w": 2}
        if isinstance(name, str):
            low = name.lower()
            if low in names:
                return names.index(low)
            if low in aliases:
                return aliases[low]
            raise ValueError(f"unknown DOF name {name!r} for Hex8Solid3D; valid names: {list(names)}")
        return int(name)

    def shape_and_derivs(self, natural_coords):
        """N (8,) and dN/d(xi,eta,zeta) (3,8) at one natural point."""
        xi, eta, zeta = natural_coords
        N = 0.125 * (1 + xi * _XI) * (1 + eta * _ETA) * (1 + zeta * _ZETA)
        dxi = 0.125 * _XI * (1 + eta * _ETA) * (1 + zeta * _ZETA)
        deta = 0.125 * _ETA * (1 + xi * _XI) * (1 + zeta * _ZETA)
        dzeta = 0.125 * _ZETA * (1 + xi * _XI) * (1 + eta * _ETA)
        return N, np.vstack([dxi, deta, dzeta])

    def stiffness_batched(self, elem_coords, D, device="cuda"):
        """(m,24,24) device tensor for coordinates (m,8,3) and D (6,6) or (m,6,6)."""
        from .operator import element_stiffness_batched
        return element_stiffness_batched(elem_coords, D, device=device)

    def stiffness(self, elem_coords, D, thickness=1.0, gauss_order=None, device="cuda"):
        """One element's 24x24 stiffness as a numpy array (computed on the GPU)."""
        if gauss_order not in (None, 2):
            raise NotImplementedError("Hex8Solid3D on the HIP backend integrates with the 2x2x2 rule only")
        ke = self.stiffness_batched(np.asarray(elem_coords, dtype=float)[None], D, device=device)
        return ke[0].cpu().numpy() * thickness


class Hex8SolidBbar(Hex8Solid3D):
    """The B-bar Hex8 (fe/elements/solids.py:225-272): at each of the 8 Gauss
    points the volumetric part of B is replaced by its element volume average,
    which relieves volumetric locking as nu -> 0.5.  Full integration only.
    On a structured grid its brick K0 keeps the reflection symmetry, so every
    matrix-free kernel and solve of the library serves it unchanged."""

    def stiffness_batched(self, elem_coords, D, device="cuda"):
        from .operator import element_stiffness_batched
        return element_stiffness_batched(elem_coords, D, device=device, bbar=True)

    def full_stiffness(self, elem_coords, D, thickness=1.0, device="cuda"):
        return self.stiffness(elem_coords, D, thickness, device=device)

    def reduced_stiffness(self, elem_coords, D, thickness=1.0):
        raise NotImplementedError("Hex8SolidBbar has no reduced-integration variant: the mean-dilatation "
                                  "correction is defined at the full 8-point rule.")
