"""Writes tests/golden/pcg_golden.npz from the reference implementation.

Run once in the build container (the reference is not available elsewhere):

    PYTHONPATH=/root/reference/fea_engine/src python tests/golden/make_golden.py

Only quantities that no closed form or invariant pins are taken from the
reference: the iteration count and residual history of its own
``preconditioned_cg`` (fe/iterative_solvers.py:362-469) driven through
``FESystem.solve_static(method="pcg")`` / ``_solve_static_iterative``
(fe/solver.py:1922-2029) on small H8 cantilevers.  The stored arrays are data
(inputs and outputs), not reference source.
"""
from __future__ import annotations

import os
import sys

import numpy as np

REF = "/root/reference/fea_engine/src"


def case(fe, nx, ny, nz, load, tol, E_elem=None):
    from fea_engine import FESystem, Hex8Solid3D, Material, D_solid3d
    from fea_engine.mesh import box_mesh
    from fea_engine import iterative_solvers as its
    from fea_engine import facet_loads
    mesh = box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = FESystem(mesh, Hex8Solid3D(), sparse=True)
    if E_elem is None:
        fs.assemble_stiffness(D_solid3d(Material(E=210e9, nu=0.3)))
    else:
        from fea_engine import coefficients
        Ds = [D_solid3d(Material(E=float(e), nu=0.3)) for e in E_elem]
        fs.assemble_stiffness(coefficients.by_element(Ds))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    if load == "facet":
        facets = []
        for conn in mesh.elements:
            for idx, _, _ in facet_loads.list_facets(Hex8Solid3D()):
                f = tuple(int(conn[i]) for i in idx)
                if np.all(np.abs(mesh.nodes[list(f), 0] - nx) < 1e-9):
                    facets.append(f)
        fs.add_consistent_facet_load(Hex8Solid3D(), facets, 2, -1e6 / (ny * nz))
    else:
        fs.add_nodal_force(mesh.nodes_on_plane(0, float(nx)), 2, -1e6)
    free = fs.free_dofs
    Kff = fs._as_solve_matrix(fs.K)[np.ix_(free, free)]
    Ff = fs.F[free]
    hist = []
    bn = np.linalg.norm(Ff)
    M = its.jacobi_preconditioner(Kff)
    x, it = its.preconditioned_cg(Kff, Ff, M=M, tol=tol,
                                  callback=lambda k, x, r: hist.append(np.linalg.norm(r) / bn))
    U = fs.solve_static(method="pcg", tol=tol)
    assert np.allclose(np.asarray(U)[free], x)
    return dict(F=np.asarray(fs.F), U=np.asarray(U), iters=it, hist=np.array(hist), free=free)


def main():
    sys.path.insert(0, REF)
    import fea_engine as fe
    out = {}
    for name, args in {
        "c1_facet_1e-10": (20, 4, 4, "facet", 1e-10, None),
        "c1_nodal_1e-8": (20, 4, 4, "nodal", 1e-8, None),
        "small_hetE_1e-8": (8, 2, 2, "facet", 1e-8, 10 ** np.random.default_rng(7).uniform(9, 12, 32)),
    }.items():
        nx, ny, nz, load, tol, E = args
        r = case(fe, nx, ny, nz, load, tol, E)
        for k, v in r.items():
            out[f"{name}/{k}"] = np.asarray(v)
        out[f"{name}/tol"] = np.asarray(tol)
        out[f"{name}/dims"] = np.asarray([nx, ny, nz])
        if E is not None:
            out[f"{name}/E_elem"] = np.asarray(E)
        print(name, "iterations", r["iters"])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
