"""Writes tests/golden/bbar_golden.npz from the reference implementation.

Run once in the build container (the reference is not available elsewhere):

    PYTHONPATH=/root/reference/fea_engine/src python tests/golden/make_bbar_golden.py

Stores the reference's Hex8SolidBbar element stiffness (fe/elements/solids.py:
225-262) on a unit cube, a distorted hex and a near-incompressible brick, and
the tip-loaded cantilever of its own B-bar test (fe/tests/test_hex8_bbar.py:
127-150: 20x4x4 hexes on 1 x 0.05 x 0.05, x = 0 clamped, -1000 split evenly
over the tip nodes, direct solve) for Hex8Solid3D and Hex8SolidBbar at
nu = 0.3 and 0.45: the full displacement and the tip ratio to Euler-Bernoulli.
The stored arrays are data, not reference source.
"""
from __future__ import annotations

import os
import sys

import numpy as np

REF = "/root/reference/fea_engine/src"
CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def main():
    sys.path.insert(0, REF)
    from fea_engine import FESystem, Hex8Solid3D, Material, D_solid3d
    from fea_engine.elements.solids import Hex8SolidBbar
    from fea_engine.mesh import box_mesh
    out = {}
    dist = CUBE + np.random.default_rng(31).uniform(-0.15, 0.15, CUBE.shape)
    for name, X, E, nu in (("cube", CUBE, 210e9, 0.3), ("distorted", dist, 1.0, 0.35),
                           ("brick_incomp", CUBE * np.array([1.0, 0.5, 0.75]), 1.0, 0.4999)):
        out[f"ke/{name}/X"] = X
        out[f"ke/{name}/E_nu"] = np.array([E, nu])
        out[f"ke/{name}/ke"] = Hex8SolidBbar().stiffness(X, D_solid3d(Material(E=E, nu=nu)))
    for cls in (Hex8Solid3D, Hex8SolidBbar):
        for nu in (0.3, 0.45):
            mesh = box_mesh(1.0, 0.05, 0.05, 20, 4, 4)
            steel = Material(E=210e9, nu=nu)
            fs = FESystem(mesh, cls())
            fs.assemble_stiffness(D_solid3d(steel))
            tip = mesh.nodes_on_plane(axis=0, value=1.0)
            fs.add_nodal_force(tip, dof_index=2, total_force=-1000.0)
            fs.fix_dofs(mesh.nodes_on_plane(axis=0, value=0.0), [0, 1, 2])
            U = np.asarray(fs.solve_static())
            tc = tip[np.argmin(np.abs(mesh.nodes[tip, 1] - 0.025))]
            I3 = 0.05 * 0.05 ** 3 / 12
            ratio = U[3 * tc + 2] / (-1000.0 / (3 * 210e9 * I3))
            key = f"cant/{cls.__name__}/{nu}"
            out[key + "/U"] = U
            out[key + "/ratio"] = np.array(ratio)
            out[key + "/tip_node"] = np.array(tc)
            print(key, "ratio", ratio)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bbar_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
