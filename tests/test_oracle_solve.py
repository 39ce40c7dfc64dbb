"""Pins the oracle's mesh, loads, assembly, Dirichlet partition, Jacobi-PCG,
SIMP sensitivities and batched solve (see the PINNING table in
oracle/h8_oracle.py)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import h8_oracle as o


def test_box_mesh_formulas_and_volume():
    nx, ny, nz = 3, 2, 4
    nodes, elems = o.box_mesh(3.0, 1.0, 2.0, nx, ny, nz)
    assert nodes.shape == ((nx + 1) * (ny + 1) * (nz + 1), 3)
    i, j, k = 2, 1, 3
    n = k * (nx + 1) * (ny + 1) + i * (ny + 1) + j
    assert np.allclose(nodes[n], [i * 1.0, j * 0.5, k * 0.5])
    e = k * nx * ny + i * ny + j
    nid = lambda a, b, c: c * (nx + 1) * (ny + 1) + a * (ny + 1) + b
    assert list(elems[e]) == [nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                              nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)]
    pts, wts = o.gauss_points(2)
    vol = sum(o.hex8_B(p, nodes[c])[1] * w for c in elems for p, w in zip(pts, wts))
    assert abs(vol - 6.0) < 1e-12


def test_facet_shares():
    rect = np.array([[5, 0, 0], [5, 2, 0], [5, 2, 3], [5, 0, 3]], dtype=float)
    assert np.allclose(o.quad4_face_shares(rect, 4.0), 4.0 * 6.0 / 4)
    skew = np.array([[0, 0, 0], [2, 0, 0], [2.5, 1, 0], [0.5, 1, 0]], dtype=float)   # parallelogram, area 2
    sh = o.quad4_face_shares(skew, 3.0)
    assert abs(sh.sum() - 6.0) < 1e-13
    c = o.cantilever(20, 4, 4)
    assert abs(c["F"].sum() + 1e6) < 1e-6
    assert abs(c["F"][2::3].sum() + 1e6) < 1e-6 and np.abs(c["F"][0::3]).max() == 0


def test_assembly_equals_dense_loop():
    """fe/tests/test_vectorized_assembly.py:62-95: the vectorised assembly equals
    the per-element loop on a distorted hex patch."""
    nodes, elems = o.box_mesh(2.0, 1.0, 1.0, 2, 1, 2)
    nodes = nodes + 0.03 * np.random.default_rng(0).uniform(-1, 1, nodes.shape)
    D = o.d_solid3d(1.0, 0.3)
    s = np.random.default_rng(1).uniform(0.5, 2.0, len(elems))
    K = o.assemble_csr(nodes, elems, D, scale=s).toarray()
    Kd = o.assemble_dense_loop(nodes, elems, D, scale=s)
    assert np.abs(K - Kd).max() < 1e-13 * np.abs(Kd).max()


def test_pcg_matches_direct_and_golden(golden):
    """fe/tests/test_iterative_solvers.py:110-116: CG agrees with the direct solve;
    plus the reference's own iteration count and history on the fixture."""
    c = o.cantilever(20, 4, 4)
    U_direct = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    hist = []
    free, fixed = o.partition(c["K"].shape[0], c["fixed_dofs"])
    bn = np.linalg.norm(c["F"][free])
    U, it = o.solve_static_pcg(c["K"], c["F"], c["fixed_dofs"], tol=1e-10,
                               callback=lambda k, x, r: hist.append(np.linalg.norm(r) / bn))
    assert np.linalg.norm(U - U_direct) / np.linalg.norm(U_direct) < 1e-8
    g = "c1_facet_1e-10"
    assert np.allclose(c["F"], golden[g + "/F"], rtol=0, atol=1e-9)
    assert it == int(golden[g + "/iters"])
    # finite-precision CG amplifies summation-order differences (the reference
    # builds K element by element, the oracle by one batched contraction): the
    # histories agree to rounding early and to a few percent at the end.
    gh = golden[g + "/hist"]
    rel = np.abs(np.array(hist) - gh) / gh
    assert rel[:30].max() < 1e-10 and rel.max() < 0.25
    assert np.linalg.norm(U - golden[g + "/U"]) / np.linalg.norm(U) < 1e-12


def test_pcg_golden_nodal_and_heterogeneous(golden):
    c = o.cantilever(20, 4, 4, load="nodal")
    U, it = o.solve_static_pcg(c["K"], c["F"], c["fixed_dofs"], tol=1e-8)
    assert it == int(golden["c1_nodal_1e-8/iters"])
    E = golden["small_hetE_1e-8/E_elem"]
    c = o.cantilever(8, 2, 2, E_elem=E)
    U, it = o.solve_static_pcg(c["K"], c["F"], c["fixed_dofs"], tol=1e-8)
    assert abs(it - int(golden["small_hetE_1e-8/iters"])) <= 1
    assert np.linalg.norm(U - golden["small_hetE_1e-8/U"]) / np.linalg.norm(U) < 1e-9


def test_pcg_errors():
    """fe/tests/test_iterative_solvers.py:118-128: an indefinite matrix raises."""
    A = np.diag([1.0, -2.0, 3.0])
    import scipy.sparse as sp
    with pytest.raises(np.linalg.LinAlgError):
        o.preconditioned_cg(sp.csr_matrix(A), np.ones(3))
    c = o.cantilever(6, 2, 2)
    with pytest.raises(RuntimeError):
        o.solve_static_pcg(c["K"], c["F"], c["fixed_dofs"], tol=1e-12, maxiter=3)
    with pytest.raises(ValueError):
        o.jacobi_preconditioner(sp.csr_matrix(np.diag([1.0, 0.0])))


def test_cantilever_locking_ratio():
    """fe/tests/test_hex8_bbar.py:127-167: tip ratio vs Euler-Bernoulli in (0.60, 0.80)."""
    E = 2.1e11
    c = o.cantilever(20, 4, 4, Lx=1.0, Ly=0.05, Lz=0.05, E=E, load="nodal", total=-1000.0)
    U = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    tip = o.nodes_on_plane(c["nodes"], 0, 1.0)
    tc = tip[np.argmin(np.abs(c["nodes"][tip, 1] - 0.025))]
    ratio = U[3 * tc + 2] / (-1000.0 / (3 * E * 0.05 * 0.05 ** 3 / 12))
    assert 0.60 < ratio < 0.80


def _simp_case(seed=3):
    nodes, elems = o.box_mesh(6.0, 2.0, 2.0, 6, 2, 2)
    D = o.d_solid3d(1.0, 0.3)
    ke = o.hex8_ke(nodes[elems[0]], D)
    rho = np.random.default_rng(seed).uniform(0.3, 1.0, len(elems))
    F = np.zeros(3 * len(nodes))
    tip = np.where((np.abs(nodes[:, 0] - 6) < 1e-9) & (np.abs(nodes[:, 1] - 1) < 1e-9) & (nodes[:, 2] < 1e-9))[0]
    F[3 * tip + 2] = -1.0
    clamp = o.nodes_on_plane(nodes, 0, 0.0)
    fixed = [3 * n + c for n in clamp for c in range(3)]
    return nodes, elems, D, ke, rho, F, fixed


def _simp_solve(nodes, elems, ke, rho, F, fixed, p=3.0, rmin=1e-9):
    s = o.simp_scale(rho, p, rmin)
    K = o.assemble_csr(nodes, elems, None, scale=s, ke=np.broadcast_to(ke, (len(elems), 24, 24)))
    return o.solve_direct(K, F, fixed), K


def test_simp_scale_grad_fd():
    """fe/tests/test_topopt.py:59-66."""
    rho = np.array([0.2, 0.5, 0.8])
    h = 1e-6
    fd = (o.simp_scale(rho + h, 3.0, 1e-3) - o.simp_scale(rho - h, 3.0, 1e-3)) / (2 * h)
    assert np.allclose(o.simp_scale_grad(rho, 3.0, 1e-3), fd, rtol=1e-8)


def test_compliance_gradient_closed_form_vs_adjoint_vs_fd():
    """fe/tests/test_topopt.py:72-122: closed form vs central differences of the
    forward solve, and vs the adjoint sensitivity."""
    nodes, elems, D, ke, rho, F, fixed = _simp_case()
    u, K = _simp_solve(nodes, elems, ke, rho, F, fixed)
    g = o.compliance_sensitivity_closed_form(ke, elems, u, rho, 3.0, 1e-9)
    ga = o.adjoint_sensitivity(K, ke, elems, fixed, u, F, rho, 3.0, 1e-9)
    assert np.allclose(g, ga, rtol=1e-9, atol=1e-14 * np.abs(g).max())
    # cancellation-free central difference: C+ - C- = -(s+ - s-) u-_e^T ke u+_e
    h = 1e-6
    for e in [0, 5, 11, 17, 23]:
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h; rm[e] -= h
        up, _ = _simp_solve(nodes, elems, ke, rp, F, fixed)
        um, _ = _simp_solve(nodes, elems, ke, rm, F, fixed)
        ds = o.simp_scale(rp[e], 3.0, 1e-9) - o.simp_scale(rm[e], 3.0, 1e-9)
        dC = -ds * o.element_quadratic(ke, elems[e:e + 1], um, up)[0]
        assert abs(dC / (2 * h) - g[e]) < 1e-6 * abs(g[e])
        # plain compliance difference agrees to the cancellation floor
        direct = (o.compliance(F, up) - o.compliance(F, um)) / (2 * h)
        assert abs(direct - g[e]) < 1e-4 * abs(g[e])


def test_batched_columns_equal_single_solves():
    """fe/tests/test_batched_solve.py:66-115: each column of a batched solve
    equals an independent solve."""
    c = o.cantilever(6, 2, 2)
    rng = np.random.default_rng(0)
    Fb = rng.normal(size=(c["K"].shape[0], 3))
    vals = {d: 1e-6 for d in c["fixed_dofs"][:6]}
    U = o.solve_static_batched(c["K"], Fb, c["fixed_dofs"], vals)
    for j in range(3):
        u1 = o.solve_direct(c["K"], Fb[:, j], c["fixed_dofs"], vals)
        assert np.linalg.norm(U[:, j] - u1) / np.linalg.norm(u1) < 1e-12


def _free_body(nx=6, ny=3, nz=3, Lx=4.0):
    nodes, elems = o.box_mesh(Lx, 1.0, 1.0, nx, ny, nz)
    K = o.assemble_csr(nodes, elems, o.d_solid3d(210e9, 0.3))
    F = np.zeros(K.shape[0])
    left, right = o.nodes_on_plane(nodes, 0, 0.0), o.nodes_on_plane(nodes, 0, Lx)
    F[3 * right] += 1000.0 / len(right)
    F[3 * left] -= 1000.0 / len(left)
    return nodes, K, F, left, right


def test_eigen_solve_free_body_is_minimum_norm():
    """fe/tests/test_singular_system.py:61-117 on a free H8 body: no Dirichlet
    DOFs, equal and opposite x-tension on the end faces.  Six rigid modes; the
    result is pinv(K) F (independent routine), balances K U = F, carries no
    mean translation and elongates the bar."""
    import scipy.linalg as sla
    nodes, K, F, left, right = _free_body()
    U, nr = o.eigen_solve(K, F)
    assert nr == 6
    Up = sla.pinv(K.toarray()) @ F
    assert np.max(np.abs(U - Up)) < 1e-8 * np.max(np.abs(Up))
    assert np.max(np.abs(K @ U - F)) < 1e-6 * np.max(np.abs(F))
    for c in range(3):
        assert abs(np.mean(U[c::3])) < 1e-9 * np.max(np.abs(U))
    assert np.mean(U[3 * right]) - np.mean(U[3 * left]) > 0


def test_eigen_solve_partial_support_and_unbalanced_load():
    """Only u_x fixed on the x=0 face: translations y, z and the rotation about x
    stay free (3 rigid modes); an x-load is balanced against them.  A y-load on
    the free body is not balanced (fe/tests/test_singular_system.py:143-162)."""
    import scipy.linalg as sla
    nodes, K, F, left, right = _free_body()
    fixed = 3 * left
    F2 = np.zeros_like(F)
    F2[3 * right] = 1000.0 / len(right)
    free, fx = o.partition(K.shape[0], fixed)
    Kff, Ff = o.reduced_system(K, F2, free, fx)
    uf, nr = o.eigen_solve(Kff, Ff)
    assert nr == 3
    up = sla.pinv(Kff.toarray()) @ Ff
    assert np.max(np.abs(uf - up)) < 1e-8 * np.max(np.abs(up))
    F3 = np.zeros_like(F)
    F3[3 * right + 1] = 1000.0 / len(right)
    with pytest.raises(np.linalg.LinAlgError, match="not self-equilibrated"):
        o.eigen_solve(K, F3)
    # a planted mistake (keeping the rigid directions instead of dropping them)
    # would leave the answer off the minimum-norm solution
    lam, V = np.linalg.eigh(K.toarray())
    bad = V @ ((V.T @ F) / np.where(np.abs(lam) < 1e-8 * lam.max(), lam.max(), lam) + np.where(
        np.abs(lam) < 1e-8 * lam.max(), 1e-3, 0.0))
    assert np.max(np.abs(bad - sla.pinv(K.toarray()) @ F)) > 1e-6


def test_element_matvec_equals_assembled_product():
    """The element-wise K u (used where a CSR K would be too large) equals the
    assembled CSR product on a heterogeneous mesh, in any chunking."""
    nx, ny, nz = 6, 3, 2
    nodes, elems = o.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    ke = o.hex8_ke(nodes[elems[0]], o.d_solid3d(1.0, 0.3))
    rng = np.random.default_rng(0)
    s = 10 ** rng.uniform(0, 3, len(elems))
    K = o.assemble_csr(nodes, elems, None, scale=s, ke=np.broadcast_to(ke, (len(elems), 24, 24)))
    u = rng.standard_normal(K.shape[0])
    ref = K @ u
    for chunk in (1, 7, 65536):
        assert np.allclose(o.element_matvec(elems, ke, s, u, chunk=chunk), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
