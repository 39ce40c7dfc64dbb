"""Writes tests/golden/constraint_golden.npz from the reference implementation.

Run once in the build container (the reference is not available elsewhere):

    PYTHONPATH=/root/reference/fea_engine/src python tests/golden/make_constraint_golden.py

Stores the reference's constrained static solve (add_constraint / tie,
fe/solver.py:1341-1397, eliminated in form_linear_system, fe/solver.py:1403-1479,
solved directly) on a 10x3x3 H8 cantilever with per-element E: the uz DOFs of
the y = 0 face tied to the y = Ly face (offset (0, Ly, 0)), a three-term
constraint with a non-zero value, a prescribed non-zero displacement, and a
consistent tip traction.  Also the reduced constraint map of a small
overlapping constraint set (fe/boundary.py:98-166).  The stored arrays are
data, not reference source.
"""
from __future__ import annotations

import os
import sys

import numpy as np

REF = "/root/reference/fea_engine/src"


def main():
    sys.path.insert(0, REF)
    from fea_engine import FESystem, Hex8Solid3D, Material, D_solid3d, coefficients
    from fea_engine.mesh import box_mesh
    from fea_engine import facet_loads
    from fea_engine.boundary import reduce_constraints
    out = {}
    nx, ny, nz = 10, 3, 3
    mesh = box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    E = 10 ** np.random.default_rng(41).uniform(9, 11, nx * ny * nz)
    fs = FESystem(mesh, Hex8Solid3D(), sparse=True)
    fs.assemble_stiffness(coefficients.by_element([D_solid3d(Material(E=float(e), nu=0.3)) for e in E]))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    mid = int(np.where((np.abs(mesh.nodes[:, 0] - 5) < 1e-9) & (np.abs(mesh.nodes[:, 1]) < 1e-9)
                       & (np.abs(mesh.nodes[:, 2] - 3) < 1e-9))[0][0])
    fs.fix_dofs([mid], [0], value=1e-4)
    y0 = mesh.nodes_on_plane(1, 0.0)
    y1 = mesh.nodes_on_plane(1, float(ny))
    y0 = np.array([n for n in y0 if mesh.nodes[n, 0] > 0.5])       # not the clamped nodes
    fs.tie(y0, y1, [2], offset=[0.0, float(ny), 0.0])
    tip = mesh.nodes_on_plane(0, float(nx))
    a, b, c = int(tip[0]), int(tip[5]), int(tip[10])
    fs.add_constraint([(a, "ux", 1.0), (b, "ux", -2.0), (c, "uy", 0.5)], value=2e-5)
    facets = []
    for conn in mesh.elements:
        for idx, _, _ in facet_loads.list_facets(Hex8Solid3D()):
            f = tuple(int(conn[i]) for i in idx)
            if np.all(np.abs(mesh.nodes[list(f), 0] - nx) < 1e-9):
                facets.append(f)
    fs.add_consistent_facet_load(Hex8Solid3D(), facets, 2, -1e6 / (ny * nz))
    U = np.asarray(fs.solve_static())
    out.update({"solve/dims": np.array([nx, ny, nz]), "solve/E": E, "solve/F": np.asarray(fs.F), "solve/U": U,
                "solve/y0": y0, "solve/y1": np.asarray(y1), "solve/mid": np.array(mid),
                "solve/abc": np.array([a, b, c])})
    # reduced map of an overlapping set (chains through shared DOFs, one prescribed DOF)
    rows = [{3: 1.0, 7: -1.0}, {7: 2.0, 11: -1.0, 5: 0.5}, {11: 1.0, 2: 1.0}, {20: 1.0, 21: -1.0}]
    vals = [0.0, 1.5, -0.25, 0.0]
    slave, mat, g = reduce_constraints(rows, vals, {2: 0.1})
    masters = sorted({m for s in slave for m in mat[int(s)]})
    M = np.zeros((len(slave), len(masters)))
    for i, s in enumerate(slave):
        for m, cf in mat[int(s)].items():
            M[i, masters.index(m)] = cf
    out.update({"reduce/slave": np.asarray(slave), "reduce/masters": np.asarray(masters), "reduce/mat": M,
                "reduce/g": np.array([g[int(s)] for s in slave])})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "constraint_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "slaves", slave, "masters", masters)


if __name__ == "__main__":
    main()
