"""GPU parity of the Hex20Solid3D path (fe/elements/solids.py:331-406) through
the C ABI: the element kernel against the oracle's 3x3x3 B^T D B sum, the
box operator and Jacobi diagonal against the oracle's assembled CSR, FESystem
solves against spsolve and the oracle's PCG iteration count, and the
reference's patch and pure-bending checks
(fe/tests/test_benchmarks_convergence.py:143-206) on the structured box:
Hex20 reproduces every linear and every quadratic equilibrium field exactly."""
import sys

import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _brick(h=(1.0, 1.0, 1.0)):
    nat = np.stack([o.XI20, o.ETA20, o.ZETA20], axis=1)
    return (nat + 1.0) * 0.5 * np.asarray(h)


def test_ke20_kernel_matches_oracle(h8, gpu):
    """Distorted elements (straight and curved edges) with per-element D and a
    shared D, against oracle.hex20_ke."""
    say("ke20 kernel")
    rng = np.random.default_rng(0)
    m = 23
    X = _brick((1.4, 0.8, 1.1))[None] + 0.08 * rng.uniform(-1, 1, (m, 20, 3))
    Ds = np.stack([o.d_solid3d(E, nu) for E, nu in zip(rng.uniform(1, 5, m), rng.uniform(0.1, 0.45, m))])
    ke = h8.Hex20Solid3D().stiffness_batched(X, Ds, device=gpu).cpu().numpy()
    ref = np.stack([o.hex20_ke(X[e], Ds[e]) for e in range(m)])
    scale = np.abs(ref).max()
    say("max |ke - oracle| / max|ke|", np.abs(ke - ref).max() / scale)
    assert np.abs(ke - ref).max() < 1e-13 * scale
    ke1 = h8.Hex20Solid3D().stiffness_batched(X[:5], Ds[0], device=gpu).cpu().numpy()
    ref1 = np.stack([o.hex20_ke(X[e], Ds[0]) for e in range(5)])
    assert np.abs(ke1 - ref1).max() < 1e-13 * np.abs(ref1).max()
    # one element through the descriptor, and the shape-function mirror against the oracle's
    assert np.allclose(h8.Hex20Solid3D().stiffness(X[0], Ds[0], device=gpu), o.hex20_ke(X[0], Ds[0]),
                       rtol=0, atol=1e-13 * scale)
    p = rng.uniform(-1, 1, 3)
    N, dN = h8.Hex20Solid3D().shape_and_derivs(p)
    N_o, dN_o = o.hex20_shape_and_derivs(*p)
    assert np.allclose(N, N_o, atol=1e-15) and np.allclose(dN, dN_o, atol=1e-15)


def _operator(h8, gpu, dims, s, h=(1.0, 0.5, 0.75), nu=0.3):
    from h8pcg.hex20 import Hex20Operator
    K0 = h8.Hex20Solid3D().stiffness_batched(_brick(h)[None], o.d_solid3d(1.0, nu), device=gpu)[0]
    return Hex20Operator(*dims, K0, s, device=gpu), K0.cpu().numpy()


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1), (3, 2, 2), (5, 4, 3), (8, 3, 5)])
def test_spmv_and_diagonal_match_oracle_csr(h8, gpu, dims):
    import torch
    say("hex20 spmv", dims)
    nx, ny, nz = dims
    h = (1.0, 0.5, 0.75)
    rng = np.random.default_rng(sum(dims))
    m = nx * ny * nz
    s = 10 ** rng.uniform(0, 2, m)
    op, K0 = _operator(h8, gpu, dims, s, h)
    mesh = h8.box_mesh_hex20(nx * h[0], ny * h[1], nz * h[2], nx, ny, nz)
    elems = np.asarray(mesh.elements)
    K = o.assemble_csr_npe(len(mesh.nodes), elems, s[:, None, None] * K0[None])
    assert op.n_dof == K.shape[0]
    for trial in range(2):
        u = rng.standard_normal(op.n_dof)
        y = op.matvec(torch.as_tensor(u, device=gpu)).cpu().numpy()
        ref = K @ u
        err = np.abs(y - ref).max() / np.abs(ref).max()
        say("  rel err", err)
        assert err < 1e-13
    # the result does not depend on the launch (fixed summation order)
    y2 = op.matvec(torch.as_tensor(u, device=gpu)).cpu().numpy()
    assert np.array_equal(y, y2)
    # Jacobi diagonal: 1/K_ii at free DOFs, 0 at fixed
    free = rng.random(op.n_dof) > 0.2
    dinv = op.set_free_mask(free.astype(np.uint8)).cpu().numpy()
    d = K.diagonal()
    assert np.all(dinv[~free] == 0.0)
    assert np.allclose(dinv[free], 1.0 / d[free], rtol=1e-13, atol=0)
    # the assembled matrix of the host mirror is the oracle's
    assert abs(op.csr(elems) - K).max() <= 1e-13 * abs(K).max()


def _cantilever(h8, nx, ny, nz, E=210e9, load=-1e6, E_elem=None, element_scale=None):
    mesh = h8.box_mesh_hex20(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex20Solid3D())
    if E_elem is None:
        fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=E, nu=0.3)), element_scale=element_scale)
    else:
        Ds = [h8.D_solid3d(h8.Material(E=float(e), nu=0.3)) for e in E_elem]
        fs.assemble_stiffness(h8.coefficients.by_element(Ds))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.add_nodal_force(mesh.nodes_on_plane(0, float(nx)), 2, load)
    return fs, mesh


@pytest.mark.parametrize("dims, hetero", [((8, 2, 2), False), ((12, 3, 3), True)])
def test_cantilever_solve_matches_spsolve_and_oracle_pcg(h8, gpu, dims, hetero):
    nx, ny, nz = dims
    say("hex20 cantilever", dims, "hetero", hetero)
    rng = np.random.default_rng(5)
    m = nx * ny * nz
    E_elem = 210e9 * 10 ** rng.uniform(-1, 0, m) if hetero else None
    fs, mesh = _cantilever(h8, nx, ny, nz, E_elem=E_elem)
    U = np.asarray(fs.solve_static(method="pcg", tol=1e-10))
    st = fs.last_solve_stats
    # the oracle's system: the 3x3x3 element matrices assembled on the host
    Es = E_elem if hetero else np.full(m, 210e9)
    ke1 = o.hex20_ke(np.asarray(mesh.nodes)[np.asarray(mesh.elements)[0]], o.d_solid3d(1.0, 0.3))
    K = o.assemble_csr_npe(len(mesh.nodes), mesh.elements, Es[:, None, None] * ke1[None])
    F = np.asarray(fs.F)
    fixed = np.array(sorted(fs.fixed_dofs))
    U_direct = o.solve_direct(K, F, fixed)
    err = np.linalg.norm(U - U_direct) / np.linalg.norm(U_direct)
    _, it_oracle = o.solve_static_pcg(K, F, fixed, tol=1e-10)
    say("  iters", st["iterations"], "oracle", it_oracle, "err vs spsolve", err)
    assert err < 1e-8
    assert abs(st["iterations"] - it_oracle) <= 2
    assert np.all(U[fixed] == 0.0)
    # the K property is the oracle's matrix
    assert abs(fs.K - K).max() <= 1e-12 * abs(K).max()
    # reactions: K U - F at the supports, balancing the applied load
    R = np.asarray(fs.reactions(U))
    R_ref = np.zeros_like(R)
    R_ref[fixed] = (K @ U_direct - F)[fixed]
    assert np.abs(R - R_ref).max() < 1e-7 * np.abs(R_ref).max()
    assert abs(R[2::3].sum() + F[2::3].sum()) < 1e-8 * abs(F[2::3].sum())


def _boundary_dirichlet_solve(h8, dims, field, nu):
    nx, ny, nz = dims
    mesh = h8.box_mesh_hex20(2.0, 1.0, 1.0, nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex20Solid3D())
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0e3, nu=nu)))
    X = np.asarray(mesh.nodes)
    exact = field(X)
    lo, hi = X.min(0), X.max(0)
    boundary = np.where(np.any((np.abs(X - lo) < 1e-9) | (np.abs(X - hi) < 1e-9), axis=1))[0]
    for n in boundary:
        for d in range(3):
            fs.fix_dofs([int(n)], [d], value=float(exact[n, d]))
    assert len(boundary) < len(X)
    U = np.asarray(fs.solve_static(tol=1e-13)).reshape(-1, 3)
    return np.abs(U - exact).max() / np.abs(exact).max()


def test_patch_test_and_pure_bending_are_reproduced_exactly(h8, gpu):
    """A linear field (the patch test) and Saint-Venant pure bending
    u = k(-xy, (x^2 + nu(y^2 - z^2))/2, nu yz), an equilibrium field with zero
    body force, prescribed on the box faces: the interior solution is the field
    itself, since both lie in the serendipity space."""
    say("hex20 patch / bending")
    G = np.array([[1e-3, 2e-4, 1e-4], [3e-4, -5e-4, 2e-4], [1e-4, 4e-4, 8e-4]])
    c = np.array([1e-3, -2e-3, 5e-4])
    err = _boundary_dirichlet_solve(h8, (4, 3, 3), lambda X: X @ G.T + c, 0.3)
    say("  patch err", err)
    assert err < 1e-10
    kappa, nu = 1e-3, 0.3

    def bending(X):
        x, y, z = X[:, 0], X[:, 1] - 0.5, X[:, 2] - 0.5
        return kappa * np.stack([-x * y, 0.5 * (x * x + nu * (y * y - z * z)), nu * y * z], axis=1)

    err = _boundary_dirichlet_solve(h8, (4, 3, 3), bending, nu)
    say("  bending err", err)
    assert err < 1e-10


def test_larger_box_converges_with_a_small_true_residual(h8, gpu):
    """A 64 x 16 x 16 Hex20 cantilever (2.2e5 DOFs) to 1e-8: the FP64
    residual ||b - K u|| / ||b|| of the returned solution, computed with the
    box operator checked above against the oracle's CSR, agrees with the
    recurrence's."""
    import torch
    nx, ny, nz = 64, 16, 16
    say("hex20 large", (nx, ny, nz))
    fs, mesh = _cantilever(h8, nx, ny, nz)
    u = fs.solve_static_device(tol=1e-8)
    st = fs.last_solve_stats
    op = fs.operator
    b = op.masked(torch.as_tensor(fs.F, device=gpu))
    r = op.masked(b, op.matvec(u))
    rel = float(torch.linalg.norm(r) / torch.linalg.norm(b))
    say("  iters", st["iterations"], "rel", st["rel_residual"], "true rel", rel, "ms", st["elapsed_ms"])
    assert st["rel_residual"] < 1e-8 and rel < 2e-8
    assert op.n_dof > 2.0e5


def test_error_paths(h8, gpu):
    say("hex20 errors")
    fs, mesh = _cantilever(h8, 4, 2, 2, E=1.0, load=-1.0)
    with pytest.raises(NotImplementedError):
        fs.solve_static(precision="mixed")
    with pytest.raises(NotImplementedError):
        fs.solve_static(precision="mixed", preconditioner="mg")
    with pytest.raises(NotImplementedError):
        fs.add_consistent_facet_load(h8.Hex20Solid3D(), [tuple(range(8))], 2, -1.0)
    with pytest.raises(NotImplementedError):
        fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), gauss_order=2)
    with pytest.raises(RuntimeError, match="did not converge"):
        fs.solve_static(tol=1e-14, maxiter=3)
    # an indefinite operator trips p^T A p <= 0; zero scales give a zero diagonal
    fs2, _ = _cantilever(h8, 4, 2, 2, E=1.0, load=-1.0, element_scale=-np.ones(16))
    with pytest.raises(np.linalg.LinAlgError):
        fs2.solve_static(method="cg")
    fs3, _ = _cantilever(h8, 4, 2, 2, E=1.0, load=-1.0, element_scale=np.zeros(16))
    with pytest.raises(ValueError, match="zero diagonal"):
        fs3.solve_static()
    # a perturbed (non-structured) Hex20 mesh is refused, not solved wrongly
    mesh4 = h8.box_mesh_hex20(4.0, 2.0, 2.0, 4, 2, 2)
    mesh4.nodes = mesh4.nodes + 1e-3 * np.random.default_rng(0).standard_normal(mesh4.nodes.shape)
    fs4 = h8.FESystem(mesh4, h8.Hex20Solid3D())
    with pytest.raises(NotImplementedError, match="structured"):
        fs4.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)))
