"""GPU parity of the multigrid preconditioner of the FP32 iterations: one
V-cycle against the oracle's restatement (scipy Galerkin products), and the
multigrid-preconditioned refined solve against the direct solve."""
import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _cantilever(h8, nx, ny, nz, E):
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda:0")
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), element_scale=E)
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.add_consistent_facet_load(h8.Hex8Solid3D(), h8.boundary_faces(mesh, 0, float(nx)), 2, -1e6 / (ny * nz))
    return fs


@pytest.mark.parametrize("dims,degree", [((32, 8, 8), 2), ((32, 8, 8), 3), ((32, 8, 8), 4), ((32, 8, 8), 6),
                                          ((40, 3, 4), 2), ((40, 3, 4), 5), ((24, 6, 6), 4)])
def test_vcycle_matches_oracle(h8, gpu, dims, degree):
    """z = V(r) on the device equals the oracle's V-cycle (FP64 arithmetic,
    coarse operators P^T A P from scipy, stored in FP32 as the library stores
    them) to FP32 accuracy, for level-0 smoothing degrees 2..6 (odd degrees
    take the other buffer order); (40, 3, 4) keeps its odd y axis on the
    coarse level; (24, 6, 6) ends on a 3 x 3 x 3 coarsest grid of 192 DOFs,
    inverted by the symmetric sweep.  (Without the FP32 storage the coarsest inverse alone
    differs by up to its condition number times the FP32 rounding.)"""
    import torch
    nx, ny, nz = dims
    rng = np.random.default_rng(5)
    E = 10 ** rng.uniform(9, 12, nx * ny * nz)
    fs = _cantilever(h8, nx, ny, nz, E)
    op = fs.operator
    op.set_mg_degree(degree)
    free = fs._free_mask().astype(bool)
    op.set_free_mask(free.astype(np.uint8))
    r = rng.standard_normal(op.n_dof) * free
    z = op.mg_vcycle(torch.as_tensor(r, dtype=torch.float32, device=op.device)).cpu().numpy().astype(float)
    c = o.cantilever(nx, ny, nz, E_elem=E)
    nodes, elems = o.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    ke = o.hex8_ke(nodes[elems[0]], o.d_solid3d(1.0, 0.3))
    Kabs = o.assemble_csr(nodes, elems, None, scale=E, ke=np.broadcast_to(np.abs(ke), (len(elems), 24, 24)))
    H = o.mg_hierarchy(c["K"].tocsr(), free, dims, K_abs=Kabs, store32=True, degree0=degree)
    zr = o.mg_vcycle(H, r.astype(np.float32).astype(float))
    err = np.linalg.norm(z - zr) / np.linalg.norm(zr)
    say("vcycle", dims, "degree", degree, "levels", [lv["A"].shape[0] for lv in H], "rel err", err)
    assert np.all(z[~free] == 0.0)
    assert err < 1e-5


@pytest.mark.parametrize("dims,span", [((20, 4, 4), 0.0), ((32, 8, 8), 3.0), ((40, 3, 4), 1.0), ((24, 6, 6), 3.0)])
def test_mg_refined_solve_matches_direct(h8, gpu, dims, span):
    """The multigrid-preconditioned refined solve (the default of
    precision='mixed') meets the FP64 true residual and agrees with the direct
    solve to 1e-8, in far fewer FP32 iterations than the Jacobi form."""
    nx, ny, nz = dims
    rng = np.random.default_rng(6)
    E = 210e9 * 10 ** rng.uniform(-span, 0, nx * ny * nz)
    fs = _cantilever(h8, nx, ny, nz, E)
    U = np.asarray(fs.solve_static(tol=1e-9, precision="mixed"))
    st = dict(fs.last_solve_stats)
    Uj = np.asarray(fs.solve_static(tol=1e-9, precision="mixed", preconditioner="jacobi"))
    stj = fs.last_solve_stats
    c = o.cantilever(nx, ny, nz, E_elem=E)
    Ud = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    free, _ = o.partition(c["K"].shape[0], c["fixed_dofs"])
    true_res = np.linalg.norm((c["F"] - c["K"] @ U)[free]) / np.linalg.norm(c["F"][free])
    err = np.linalg.norm(U - Ud) / np.linalg.norm(Ud)
    errj = np.linalg.norm(Uj - Ud) / np.linalg.norm(Ud)
    say("mg", dims, "fp32 iters", st["inner_iterations"], "replacements", st["outer"], "true_res", true_res,
        "reported", st["true_rel_residual"],
        "err", err, "| jacobi iters", stj["inner_iterations"], "err", errj, "ms", st["elapsed_ms"],
        stj["elapsed_ms"])
    assert st["preconditioner"] == "mg" and stj["preconditioner"] == "jacobi"
    assert true_res < 1e-9
    # the FP64 residual of a 1e-9 solution carries about 1e-16 cond(K) of
    # evaluation noise (the device and the CPU sum in different orders)
    assert abs(true_res - st["true_rel_residual"]) < 2e-2 * true_res
    assert err < 1e-8
    assert st["inner_iterations"] < 120
    assert st["inner_iterations"] < stj["inner_iterations"] / 3


def test_mg_degree_solves_agree(h8, gpu):
    """The refined solve at level-0 smoothing degrees 2..6 and the library's
    choice (0) reaches the direct solution to 1e-8; a stronger smoother takes
    fewer FP32 iterations."""
    nx, ny, nz = 32, 8, 8
    E = 210e9 * 10 ** np.random.default_rng(6).uniform(-3, 0, nx * ny * nz)
    fs = _cantilever(h8, nx, ny, nz, E)
    c = o.cantilever(nx, ny, nz, E_elem=E)
    Ud = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    its = {}
    for deg in (2, 3, 4, 6, 0):
        fs.operator.set_mg_degree(deg)
        U = np.asarray(fs.solve_static(tol=1e-9, precision="mixed"))
        st = fs.last_solve_stats
        err = np.linalg.norm(U - Ud) / np.linalg.norm(Ud)
        say("degree", deg, "fp32 iters", st["inner_iterations"], "replacements", st["outer"], "err", err)
        assert st["preconditioner"] == "mg"
        assert err < 1e-8
        its[deg] = st["inner_iterations"]
    assert its[6] < its[2] and its[4] < its[2]
    with pytest.raises(ValueError):
        fs.operator.set_mg_degree(7)


def test_mg_deterministic_and_warm_start(h8, gpu):
    """Two solves give bitwise equal results; a warm start from the solution
    returns after the first FP64 residual with no FP32 iteration."""
    nx, ny, nz = 32, 8, 8
    E = 10 ** np.random.default_rng(8).uniform(9, 11, nx * ny * nz)
    fs = _cantilever(h8, nx, ny, nz, E)
    U1 = np.asarray(fs.solve_static(tol=1e-9, precision="mixed"))
    U2 = np.asarray(fs.solve_static(tol=1e-9, precision="mixed"))
    assert np.array_equal(U1, U2)
    fs.solve_static(tol=1e-9, precision="mixed", x0=U1)
    say("warm start", fs.last_solve_stats["inner_iterations"], fs.last_solve_stats["preconditioner"])
    assert fs.last_solve_stats["inner_iterations"] == 0
    assert fs.last_solve_stats["preconditioner"] == "mg"


def test_mg_free_body_uses_the_jacobi_form(h8, gpu):
    """Supports that leave rigid-body motions free make every coarse operator
    singular too: a free-body mixed solve runs the Jacobi form and returns the
    minimum-norm solution, as the FP64 solve does."""
    nx, ny, nz = 16, 4, 4
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda:0")
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)))
    # self-equilibrated pair of forces along x at the two ends
    F = np.zeros(fs.n_dof)
    F[3 * mesh.nodes_on_plane(0, 0.0)] = -1.0
    F[3 * mesh.nodes_on_plane(0, float(nx))] = 1.0
    fs.F[:] = F
    U = np.asarray(fs.solve_static(tol=1e-9, precision="mixed"))
    assert fs.last_solve_stats["preconditioner"] == "jacobi"
    U64 = np.asarray(fs.solve_static(tol=1e-10))
    assert np.linalg.norm(U - U64) / np.linalg.norm(U64) < 1e-7


def _roller_box(h8, nx, ny, nz, L, E, prescribed=None):
    """Non-cubic bricks (L / n per axis) on symmetry rollers: u_x = 0 on x = 0,
    u_y = 0 on y = 0, u_z = 0 on z = 0 (single-component masks, no rigid-body
    motion left), a -z load on the top face's far edge, and optionally u_x =
    `prescribed` on the x = L face."""
    mesh = h8.box_mesh(*L, nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda:0")
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), element_scale=E)
    for axis in range(3):
        fs.fix_dofs(mesh.nodes_on_plane(axis, 0.0), [axis])
    if prescribed is not None:
        fs.fix_dofs(mesh.nodes_on_plane(0, L[0]), [0], value=prescribed)
    top = np.intersect1d(mesh.nodes_on_plane(2, L[2]), mesh.nodes_on_plane(0, L[0]))
    fs.add_nodal_force(top, 2, -1e3)
    return fs, mesh


def test_vcycle_matches_oracle_non_cubic_rollers(h8, gpu):
    """The V-cycle on non-cubic bricks (0.25 x 0.2 x 0.4) with single-component
    roller masks equals the oracle's (its K assembled from the real geometry)."""
    import torch
    nx, ny, nz = 24, 6, 10
    L = (6.0, 1.2, 4.0)
    rng = np.random.default_rng(11)
    E = 10 ** rng.uniform(9, 11, nx * ny * nz)
    fs, mesh = _roller_box(h8, nx, ny, nz, L, E)
    op = fs.operator
    free = fs._free_mask().astype(bool)
    op.set_free_mask(free.astype(np.uint8))
    r = rng.standard_normal(op.n_dof) * free
    z = op.mg_vcycle(torch.as_tensor(r, dtype=torch.float32, device=op.device)).cpu().numpy().astype(float)
    nodes, elems = o.box_mesh(*L, nx, ny, nz)
    D = o.d_solid3d(1.0, 0.3)
    K = o.assemble_csr(nodes, elems, D, scale=E)
    ke = o.hex8_ke(nodes[elems[0]], D)
    Kabs = o.assemble_csr(nodes, elems, None, scale=E, ke=np.broadcast_to(np.abs(ke), (len(elems), 24, 24)))
    H = o.mg_hierarchy(K.tocsr(), free, (nx, ny, nz), K_abs=Kabs, store32=True, degree0=4)
    zr = o.mg_vcycle(H, r.astype(np.float32).astype(float))
    err = np.linalg.norm(z - zr) / np.linalg.norm(zr)
    say("vcycle rollers", (nx, ny, nz), "levels", [lv["A"].shape[0] for lv in H], "rel err", err)
    assert np.all(z[~free] == 0.0)
    assert err < 1e-5


@pytest.mark.parametrize("prescribed", [None, 1e-4])
def test_mg_refined_solve_rollers(h8, gpu, prescribed):
    """The multigrid-refined solve on non-cubic bricks with roller supports (and
    a prescribed face displacement) agrees with the oracle's direct solve to
    1e-8, with the FP64 true residual below the tolerance."""
    nx, ny, nz = 24, 6, 10
    L = (6.0, 1.2, 4.0)
    E = 10 ** np.random.default_rng(12).uniform(9, 11, nx * ny * nz)
    fs, mesh = _roller_box(h8, nx, ny, nz, L, E, prescribed)
    U = np.asarray(fs.solve_static(tol=1e-9, precision="mixed"))
    st = dict(fs.last_solve_stats)
    nodes, elems = o.box_mesh(*L, nx, ny, nz)
    K = o.assemble_csr(nodes, elems, o.d_solid3d(1.0, 0.3), scale=E)
    F = np.asarray(fs.F, dtype=float)
    fixed = sorted(fs.fixed_dofs)
    vals = {d: fs.fixed_dof_values[d] for d in fixed} if prescribed is not None else None
    Ud = o.solve_direct(K, F, fixed, vals)
    free, _ = o.partition(K.shape[0], fixed)
    rhs = F - K @ np.where(np.isin(np.arange(K.shape[0]), fixed), Ud, 0.0)
    true_res = np.linalg.norm((F - K @ U)[free]) / np.linalg.norm(rhs[free])
    err = np.linalg.norm(U - Ud) / np.linalg.norm(Ud)
    say("rollers", prescribed, "fp32 iters", st["inner_iterations"], "replacements", st["outer"], "err", err,
        "true_res", true_res)
    assert st["preconditioner"] == "mg"
    assert np.allclose(U[fixed], Ud[fixed], rtol=0, atol=1e-15)
    assert err < 1e-8
    assert true_res < 1e-9 * 1.01
