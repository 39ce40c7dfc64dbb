"""Writes tests/golden/recovery_golden.npz from the reference implementation.

Run once in the build container (the reference is not available elsewhere):

    PYTHONPATH=/root/reference/fea_engine/src python tests/golden/make_recovery_golden.py

Stores the inputs (mesh dims and lengths, per-element Young's modulus, a seeded
displacement vector) and the outputs of the reference's stress / strain / von
Mises recovery (fe/recovery.py:258-420 through FESystem.stress / strain /
von_mises, fe/solver.py:413-430): the lumped projection of the 2x2x2
Gauss-point values to the nodes and the element volume averages.  Case "het"
has per-element material (coefficients.by_element: the reference's
per-element loop, fe/recovery.py:234-255), case "const" one D (its batched
path, fe/recovery.py:193-231).  The stored arrays are data, not reference source.
"""
from __future__ import annotations

import os
import sys

import numpy as np

REF = "/root/reference/fea_engine/src"


def case(dims, lengths, E_elem, seed):
    from fea_engine import FESystem, Hex8Solid3D, Material, D_solid3d, coefficients
    from fea_engine.mesh import box_mesh
    nx, ny, nz = dims
    mesh = box_mesh(*lengths, nx, ny, nz)
    fs = FESystem(mesh, Hex8Solid3D(), sparse=True)
    if np.ndim(E_elem) == 0:
        fs.assemble_stiffness(D_solid3d(Material(E=float(E_elem), nu=0.3)))
    else:
        fs.assemble_stiffness(coefficients.by_element([D_solid3d(Material(E=float(e), nu=0.3)) for e in E_elem]))
    U = np.random.default_rng(seed).normal(size=fs.n_dof) * 1e-4
    out = dict(dims=np.asarray(dims), lengths=np.asarray(lengths), E=np.asarray(E_elem, dtype=float), U=U)
    out["stress_nodes"] = np.asarray(fs.stress(U, at="nodes"))
    out["stress_elements"] = np.asarray(fs.stress(U, at="elements"))
    out["strain_nodes"] = np.asarray(fs.strain(U, at="nodes"))
    out["strain_elements"] = np.asarray(fs.strain(U, at="elements"))
    out["vm_nodes"] = np.asarray(fs.von_mises(U, at="nodes"))
    out["vm_elements"] = np.asarray(fs.von_mises(U, at="elements"))
    return out


def main():
    sys.path.insert(0, REF)
    res = {}
    E_het = 10 ** np.random.default_rng(13).uniform(9, 11, 4 * 3 * 2)
    for name, args in {"het": ((4, 3, 2), (4.0, 1.5, 1.5), E_het, 14),
                       "const": ((3, 2, 2), (3.0, 2.0, 2.0), 210e9, 15)}.items():
        for k, v in case(*args).items():
            res[f"{name}/{k}"] = np.asarray(v)
        print(name, {k: np.shape(v) for k, v in res.items() if k.startswith(name)})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "recovery_golden.npz")
    np.savez_compressed(path, **res)
    print("wrote", path)


if __name__ == "__main__":
    main()
