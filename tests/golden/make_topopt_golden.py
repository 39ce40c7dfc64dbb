"""Writes tests/golden/topopt_golden.npz from the reference implementation.

Run once in the build container (the reference is not available elsewhere):

    PYTHONPATH=/root/reference/fea_engine/src python tests/golden/make_topopt_golden.py

Stores inputs and outputs of the reference's SIMP loop pieces
(fe/topopt.py:235-377): the sensitivity filter (build_filter_weights +
apply_sensitivity_filter) on a 6x4x3 grid of 1 x 0.5 x 0.75 bricks, one OC
update (oc_update) of seeded densities and sensitivities, and a 6-iteration
filtered SIMP run (topology_optimize_compliance, closed-form sensitivities) on
an 8x4x2 H8 cantilever.  The stored arrays are data, not reference source.
"""
from __future__ import annotations

import os
import sys

import numpy as np

REF = "/root/reference/fea_engine/src"


def main():
    sys.path.insert(0, REF)
    from fea_engine import FESystem, Hex8Solid3D, Material, D_solid3d
    from fea_engine.mesh import box_mesh
    from fea_engine import topopt
    out = {}
    rng = np.random.default_rng(21)
    # filter
    mesh = box_mesh(6.0, 2.0, 2.25, 6, 4, 3)
    conn = np.asarray(mesh.elements)
    cent = np.asarray(mesh.nodes)[conn].mean(axis=1)
    rho = rng.uniform(0.2, 1.0, len(conn))
    dc = -rng.uniform(0.1, 5.0, len(conn))
    H = topopt.build_filter_weights(cent, 1.3)
    out.update({"filter/dims": np.array([6, 4, 3]), "filter/h": np.array([1.0, 0.5, 0.75]),
                "filter/radius": np.array(1.3), "filter/rho": rho, "filter/dc": dc,
                "filter/out": topopt.apply_sensitivity_filter(H, rho, dc)})
    # OC update
    rho = rng.uniform(0.2, 0.8, 200)
    dc = -10 ** rng.uniform(-2, 2, 200)
    dc[7] = 0.5                                        # a positive entry (clipped to -1e-12)
    out.update({"oc/rho": rho, "oc/dc": dc, "oc/vf": np.array(0.45), "oc/move": np.array(0.2),
                "oc/lo": np.array(1e-3), "oc/out": topopt.oc_update(rho, dc, 0.45, move=0.2, rho_min_bound=1e-3)})
    # SIMP run
    nx, ny, nz = 8, 4, 2
    mesh = box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = FESystem(mesh, Hex8Solid3D())
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    nodes = np.asarray(mesh.nodes)
    tip = int(np.where((np.abs(nodes[:, 0] - nx) < 1e-9) & (np.abs(nodes[:, 1] - ny / 2) < 1e-9)
                       & (np.abs(nodes[:, 2]) < 1e-9))[0][0])
    F = np.zeros(fs.n_dof)
    F[3 * tip + 2] = -1.0
    D = D_solid3d(Material(E=1.0, nu=0.3))
    res = topopt.topology_optimize_compliance(fs, D, F, volume_fraction=0.4, p=3.0, rho_min=1e-2, move=0.2,
                                              n_iter=6, filter_radius=1.5, sensitivity="closed_form", tol=1e-6)
    out.update({"run/dims": np.array([nx, ny, nz]), "run/F": F, "run/fixed": np.asarray(fs.fixed_dofs_array),
                "run/rho": res["rho"], "run/hist": np.asarray(res["compliance_history"]), "run/u": res["u_final"]})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "topopt_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "history", out["run/hist"])


if __name__ == "__main__":
    main()
