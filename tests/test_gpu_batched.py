"""GPU parity of the batched solves: every (instance, load case) system against
the oracle's direct solve, per-system convergence masks, and the multi-RHS
FESystem mirror of solve_static_batched."""
import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu


def test_solve_static_batched_matches_oracle(gpu):
    import h8pcg
    from h8pcg.batched import solve_static_batched
    nx, ny, nz = 10, 3, 3
    mesh = h8pcg.box_mesh(10.0, 3.0, 3.0, nx, ny, nz)
    fs = h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D())
    E = 10 ** np.random.default_rng(0).uniform(9, 10, nx * ny * nz)
    fs.assemble_stiffness(h8pcg.D_solid3d(h8pcg.Material(E=1.0, nu=0.3)), element_scale=E)
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.fix_dofs(mesh.nodes_on_plane(0, 10.0), [0], value=1e-4)
    rng = np.random.default_rng(1)
    Fb = rng.normal(size=(fs.n_dof, 5)) * 1e3
    Fb[:, 2] = 0.0                           # a column driven by the prescribed values only
    U, st = solve_static_batched(fs, Fb, tol=1e-11, return_stats=True)
    print("[gpu-test] batched stats", st["iterations"], st["elapsed_ms"], flush=True)
    c = o.cantilever(nx, ny, nz, E_elem=E, load="nodal", total=0.0)
    fixed = sorted(fs.fixed_dofs)
    vals = {d: fs.fixed_dof_values[d] for d in fixed}
    Uref = o.solve_static_batched(c["K"], Fb, fixed, vals)
    for j in range(5):
        err = np.linalg.norm(U[:, j] - Uref[:, j]) / np.linalg.norm(Uref[:, j])
        assert err < 1e-8, (j, err)
    assert np.all(st["status"] == 0)


def test_uneven_instances_and_masks(gpu):
    import torch
    import h8pcg
    from h8pcg.batched import solve_many
    from h8pcg.workloads import cantilever, tip_traction_loads
    rng = np.random.default_rng(7)
    systems = []
    refs = []
    for nx, k in [(8, 2), (12, 4), (16, 1), (20, 3), (8, 5)]:
        ny = nz = nx // 4
        E = 10 ** rng.uniform(9, 11, nx * ny * nz)
        fs = cantilever(nx, ny, nz, E_elem=E, load="none")
        F = tip_traction_loads(fs, nx, ny, nz, k, rng)
        if nx == 16:
            F[:] = 0.0                        # zero load: retired at once (0 iterations)
        systems.append((fs, F))
        c = o.cantilever(nx, ny, nz, E_elem=E, load="nodal", total=0.0)
        refs.append((c, F))
    X, st = solve_many(systems, tol=1e-10)
    print("[gpu-test] solve_many", st["iterations"].tolist(), st["elapsed_ms"], st["launches"], flush=True)
    assert np.all(st["status"] == 0)
    s = 0
    for (c, F), x in zip(refs, X):
        x = x.cpu().numpy()
        for j in range(F.shape[1]):
            if not np.any(F[:, j]):
                assert st["iterations"][s] == 0 and np.all(x[:, j] == 0)
            else:
                ref = o.solve_direct(c["K"], F[:, j], c["fixed_dofs"])
                err = np.linalg.norm(x[:, j] - ref) / np.linalg.norm(ref)
                assert err < 1e-7, (s, err)
            s += 1
    # the batched recurrence is the single-system one: same iteration counts (+-2)
    fs, F = systems[1]
    fs.F = F[:, 0].copy()
    fs.solve_static(tol=1e-10)
    assert abs(fs.last_solve_stats["iterations"] - st["iterations"][2]) <= 2
