"""Isotropic linear-elastic material (fe/material.py:25-37, 650-659)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Material:
    E: float
    nu: float
    rho: float = 0.0

    @property
    def G(self):
        return self.E / (2.0 * (1.0 + self.nu))


def D_solid3d(mat: Material) -> np.ndarray:
    """6x6 isotropic constitutive matrix, Voigt order [xx, yy, zz, xy, yz, zx]
    with engineering shear strains (diagonal shear term G)."""
    E, nu = mat.E, mat.nu
    c = E / ((1 + nu) * (1 - 2 * nu))
    g = (1 - 2 * nu) / 2
    return c * np.array([
        [1 - nu, nu, nu, 0, 0, 0],
        [nu, 1 - nu, nu, 0, 0, 0],
        [nu, nu, 1 - nu, 0, 0, 0],
        [0, 0, 0, g, 0, 0],
        [0, 0, 0, 0, g, 0],
        [0, 0, 0, 0, 0, g]])
