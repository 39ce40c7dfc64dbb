"""GPU parity of the constrained solve (h8_pcg_constrained_f64 through
FESystem.add_constraint / tie / solve_static) against the reference's own
direct solve on the committed fixture and against the oracle on a face tie."""
import os
import sys

import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "constraint_golden.npz")


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def h8(gpu):
    import h8pcg
    return h8pcg


def test_constrained_solve_matches_reference(h8):
    g = np.load(GOLD)
    nx, ny, nz = (int(v) for v in g["solve/dims"])
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
    fs.assemble_stiffness(h8.coefficients.by_element([h8.D_solid3d(h8.Material(E=float(e), nu=0.3))
                                                      for e in g["solve/E"]]))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.fix_dofs([int(g["solve/mid"])], [0], value=1e-4)
    fs.tie(g["solve/y0"], g["solve/y1"], [2], offset=[0.0, float(ny), 0.0])
    a, b, c = (int(v) for v in g["solve/abc"])
    fs.add_constraint([(a, "ux", 1.0), (b, "ux", -2.0), (c, "uy", 0.5)], value=2e-5)
    fs.add_consistent_facet_load(h8.Hex8Solid3D(), h8.boundary_faces(mesh, 0, float(nx)), 2, -1e6 / (ny * nz))
    assert np.allclose(fs.F, g["solve/F"], rtol=0, atol=1e-6)
    U = np.asarray(fs.solve_static(tol=1e-12))
    ref = g["solve/U"]
    err = np.linalg.norm(U - ref) / np.linalg.norm(ref)
    say("constrained vs reference", err, fs.last_solve_stats)
    assert err < 1e-9
    assert abs(U[3 * a] - 2 * U[3 * b] + 0.5 * U[3 * c + 1] - 2e-5) < 1e-12
    assert U[3 * int(g["solve/mid"])] == 1e-4
    with pytest.raises(NotImplementedError):
        fs.solve_static(precision="mixed")


def test_face_tie_matches_oracle(h8):
    nx, ny, nz = 24, 6, 6
    rng = np.random.default_rng(17)
    E = 10 ** rng.uniform(9, 11, nx * ny * nz)
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
    fs.assemble_stiffness(h8.coefficients.by_element([h8.D_solid3d(h8.Material(E=float(e), nu=0.3)) for e in E]))
    clamp = mesh.nodes_on_plane(0, 0.0)
    fs.fix_dofs(clamp, [0, 1, 2])
    tip = mesh.nodes_on_plane(0, float(nx))
    centre = int(tip[np.argmin(np.linalg.norm(np.asarray(mesh.nodes)[tip] - [nx, ny / 2, nz / 2], axis=1))])
    # the tip face moves rigidly in z with its centre node (a star of ties), and the
    # z = 0 and z = Lz faces are tied in x (periodic-style pairs)
    for n in tip:
        if int(n) != centre:
            fs.add_constraint([(int(n), "uz", 1.0), (centre, "uz", -1.0)])
    z0 = np.array([n for n in mesh.nodes_on_plane(2, 0.0) if mesh.nodes[n, 0] > 0.5 and mesh.nodes[n, 0] < nx - 0.5])
    fs.tie(z0, mesh.nodes_on_plane(2, float(nz)), [0], offset=[0.0, 0.0, float(nz)])
    fs.add_nodal_force([centre], 2, -1e6)
    say("face tie", len(fs._constraints), "constraints")
    U = np.asarray(fs.solve_static(tol=1e-12))
    c = o.cantilever(nx, ny, nz, E_elem=E)
    cons = [(d, co, v) for d, co, v in fs._constraints]
    Uo = o.solve_constrained(c["K"], np.asarray(fs.F), c["fixed_dofs"], np.zeros(len(c["fixed_dofs"])), cons)
    err = np.linalg.norm(U - Uo) / np.linalg.norm(Uo)
    say("face tie vs oracle", err, fs.last_solve_stats["iterations"])
    assert err < 1e-9
    for d, co, v in cons:
        assert abs(co @ U[d] - v) < 1e-12 * np.abs(U).max()


def test_form_linear_system_bring_your_own_solver(h8):
    """form_linear_system: the reduced operator (products on the GPU) and load
    equal the oracle's T^T K T and T^T (F - K w); scipy's CG on them, recovered,
    is the reference's constrained solution."""
    import scipy.sparse.linalg as spla
    g = np.load(GOLD)
    nx, ny, nz = (int(v) for v in g["solve/dims"])
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
    fs.assemble_stiffness(h8.coefficients.by_element([h8.D_solid3d(h8.Material(E=float(e), nu=0.3))
                                                      for e in g["solve/E"]]))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.fix_dofs([int(g["solve/mid"])], [0], value=1e-4)
    fs.tie(g["solve/y0"], g["solve/y1"], [2], offset=[0.0, float(ny), 0.0])
    a, b, c = (int(v) for v in g["solve/abc"])
    fs.add_constraint([(a, "ux", 1.0), (b, "ux", -2.0), (c, "uy", 0.5)], value=2e-5)
    fs.add_consistent_facet_load(h8.Hex8Solid3D(), h8.boundary_faces(mesh, 0, float(nx)), 2, -1e6 / (ny * nz))
    rs = fs.form_linear_system()
    # the oracle's reduced system with the same elimination
    co = o.cantilever(nx, ny, nz, E_elem=g["solve/E"])
    fixed = fs.fixed_dofs_array
    fv = dict(zip(fixed.tolist(), fs._fixed_dof_values_array(fixed).tolist()))
    rows = [dict(zip(d.tolist(), cc.tolist())) for d, cc, _ in fs._constraints]
    slave, mat, gg = o.reduce_constraints(rows, [v for _, _, v in fs._constraints], fv)
    keep = np.setdiff1d(np.setdiff1d(np.arange(fs.n_dof), fixed), slave)
    assert np.array_equal(rs.free, keep)
    T = np.zeros((fs.n_dof, len(keep)))
    T[keep, np.arange(len(keep))] = 1.0
    kpos = {int(d): i for i, d in enumerate(keep)}
    for s in slave:
        for m, cf in mat[int(s)].items():
            T[int(s), kpos[int(m)]] = -cf
    w = np.zeros(fs.n_dof)
    w[fixed] = [fv[int(d)] for d in fixed]
    w[slave] = [gg[int(s)] for s in slave]
    K = co["K"].toarray()
    Kr = T.T @ K @ T
    Fr = T.T @ (fs.F - K @ w)
    v = np.random.default_rng(3).normal(size=len(keep))
    say("form_linear_system", len(keep), np.abs(rs.K @ v - Kr @ v).max() / np.abs(Kr @ v).max(),
        np.abs(rs.F - Fr).max() / np.abs(Fr).max())
    assert np.abs(rs.K @ v - Kr @ v).max() <= 1e-12 * np.abs(Kr @ v).max()
    assert np.abs(rs.F - Fr).max() <= 1e-12 * np.abs(Fr).max()
    x, info = spla.cg(rs.K, rs.F, rtol=1e-13, maxiter=20000)
    U = np.asarray(rs.recover(x))
    ref = g["solve/U"]
    assert info == 0 and np.linalg.norm(U - ref) / np.linalg.norm(ref) < 1e-8
    assert np.linalg.norm(np.asarray(rs.solve(tol=1e-12)) - ref) / np.linalg.norm(ref) < 1e-9
