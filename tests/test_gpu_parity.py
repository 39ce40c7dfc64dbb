"""GPU parity: libh8pcg.so (through the C ABI, via h8pcg) against the oracle."""
import sys

import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def h8(gpu):
    import h8pcg
    from h8pcg import _lib
    buf = __import__("ctypes").create_string_buffer(256)
    _lib.check(_lib.lib().h8_device_info(buf, 256))
    say("device", buf.value.decode())
    return h8pcg


def test_ke_kernel_matches_oracle(h8, gpu):
    """The batched element kernel (J, detJ and the physical shape gradients in
    registers, fe/vectorized_assembly.py:94-214) against the oracle's per-element
    B^T D B sum on distorted hexes: the role of fe/tests/test_mesh_transform.py:
    67-85 and fe/tests/test_torch_vectorized_assembly.py:127-159 (batched shape
    gradients and stiffness equal the per-element ones)."""
    say("ke kernel")
    rng = np.random.default_rng(0)
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    X = base[None] * np.array([1.5, 0.7, 1.1]) + 0.12 * rng.uniform(-1, 1, (37, 8, 3))
    Ds = np.stack([o.d_solid3d(E, nu) for E, nu in zip(rng.uniform(1, 5, 37), rng.uniform(0.1, 0.45, 37))])
    ke = h8.Hex8Solid3D().stiffness_batched(X, Ds, device=gpu).cpu().numpy()
    ref = o.hex8_ke_batched(X, Ds)
    assert np.abs(ke - ref).max() < 1e-13 * np.abs(ref).max()
    ke1 = h8.Hex8Solid3D().stiffness_batched(X, Ds[0], device=gpu).cpu().numpy()
    assert np.abs(ke1 - o.hex8_ke_batched(X, Ds[0])).max() < 1e-13 * np.abs(ref).max()


def _op(h8, gpu, nx, ny, nz, s, h=(1.0, 1.0, 1.0), nu=0.3):
    from h8pcg.operator import GridOperator
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float) * h
    K0 = h8.Hex8Solid3D().stiffness_batched(X[None], o.d_solid3d(1.0, nu), device=gpu)[0]
    return GridOperator(nx, ny, nz, K0, s, device=gpu)


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 3, 2), (13, 9, 7), (20, 4, 4), (33, 17, 9)])
def test_spmv_f64_f32_match_oracle_csr(h8, gpu, dims):
    import torch
    nx, ny, nz = dims
    say("spmv", dims)
    rng = np.random.default_rng(sum(dims))
    m = nx * ny * nz
    s = 10 ** rng.uniform(0, 2, m)
    op = _op(h8, gpu, nx, ny, nz, s, h=(1.0, 0.5, 0.75))
    nodes, elems = o.box_mesh(nx * 1.0, ny * 0.5, nz * 0.75, nx, ny, nz)
    K = o.assemble_csr(nodes, elems, o.d_solid3d(1.0, 0.3), scale=s)
    u = rng.normal(size=K.shape[0])
    ref = K @ u
    # the FP32 product of the FP32-rounded input: FP64 element products leave
    # only the rounding of the outputs and of the 8-term node sums
    ref_u32 = K @ u.astype(np.float32).astype(np.float64)
    for variant in (1, 2, 3, 4, 5, 6):  # node, element MFMA, persistent, two per node, sectors, narrow sectors
        op.set_variant(variant)
        y = op.matvec(torch.as_tensor(u, device=gpu)).cpu().numpy()
        e64 = np.abs(y - ref).max() / np.abs(ref).max()
        for arith in ((32, 64) if variant in (5, 6) else (0,)):
            op.set_arith(arith)
            y32 = op.matvec32(torch.as_tensor(u, device=gpu)).cpu().numpy()
            e32 = np.abs(y32 - ref).max() / np.abs(ref).max()
            e32r = np.abs(y32 - ref_u32).max() / np.abs(ref).max()
            say("variant", variant, "arith", arith, "err64", e64, "err32", e32, "vs K u32", e32r)
            assert e32 <= 2e-5
            if arith == 64:
                assert e32r <= 1e-6
        op.set_arith(0)
        assert e64 <= 1e-12


def _fesys(h8, nx, ny, nz, E=210e9, E_elem=None, load="facet"):
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
    if E_elem is None:
        fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=E, nu=0.3)))
    else:
        Ds = [h8.D_solid3d(h8.Material(E=float(e), nu=0.3)) for e in E_elem]
        fs.assemble_stiffness(h8.coefficients.by_element(Ds))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    if load == "facet":
        fs.add_consistent_facet_load(h8.Hex8Solid3D(), h8.boundary_faces(mesh, 0, float(nx)), 2, -1e6 / (ny * nz))
    else:
        fs.add_nodal_force(mesh.nodes_on_plane(0, float(nx)), 2, -1e6)
    return fs


def test_config1_pcg_matches_reference_and_spsolve(h8, gpu, golden):
    say("config1 pcg")
    fs = _fesys(h8, 20, 4, 4)
    g = "c1_facet_1e-10"
    assert np.allclose(fs.F, golden[g + "/F"], rtol=0, atol=1e-9)
    U = fs.solve_static(method="pcg", tol=1e-10)
    st = fs.last_solve_stats
    c = o.cantilever(20, 4, 4)
    U_direct = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    err = np.linalg.norm(np.asarray(U) - U_direct) / np.linalg.norm(U_direct)
    say("iters", st["iterations"], "golden", int(golden[g + "/iters"]), "err vs spsolve", err)
    assert err < 1e-8
    assert abs(st["iterations"] - int(golden[g + "/iters"])) <= 2
    assert np.linalg.norm(np.asarray(U) - golden[g + "/U"]) / np.linalg.norm(U_direct) < 1e-8
    # fixed DOFs exactly zero
    assert np.all(np.asarray(U)[fs.fixed_dofs_array] == 0.0)


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6, 7])
def test_pcg_variants_agree(h8, gpu, variant):
    say("pcg variant", variant)
    rng = np.random.default_rng(4)
    nx, ny, nz = 16, 5, 4
    E = 10 ** rng.uniform(9, 11, nx * ny * nz)
    fs = _fesys(h8, nx, ny, nz, E_elem=E)
    fs.operator.set_variant(variant)
    c = o.cantilever(nx, ny, nz, E_elem=E)
    Ud = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    for prec in ("fp64", "mixed"):
        U = np.asarray(fs.solve_static(tol=1e-9, precision=prec))
        err = np.linalg.norm(U - Ud) / np.linalg.norm(Ud)
        say(prec, fs.last_solve_stats["iterations"], err)
        assert err < 1e-8


def test_resident_solver_golden_and_streaming_agreement(h8, gpu, golden):
    """The one-launch resident solver (variant 7): the reference's iteration
    count on the 20x4x4 cantilever, the spsolve solution, and agreement with the
    streaming kernels on a grid split over many blocks."""
    from h8pcg import _lib
    say("resident config1")
    fs = _fesys(h8, 20, 4, 4)
    fs.operator.set_variant(_lib.VARIANT_RESIDENT)
    U = np.asarray(fs.solve_static(method="pcg", tol=1e-10))
    st = fs.last_solve_stats
    c = o.cantilever(20, 4, 4)
    U_direct = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    err = np.linalg.norm(U - U_direct) / np.linalg.norm(U_direct)
    say("resident iters", st["iterations"], "launches", st["launches"], "err", err)
    assert err < 1e-8
    assert abs(st["iterations"] - int(golden["c1_facet_1e-10/iters"])) <= 2
    assert st["launches"] == 1
    assert np.all(U[fs.fixed_dofs_array] == 0.0)
    rng = np.random.default_rng(7)
    nx, ny, nz = 40, 10, 9
    E = 10 ** rng.uniform(9, 11, nx * ny * nz)
    fs = _fesys(h8, nx, ny, nz, E_elem=E)
    res = {}
    for variant in (_lib.VARIANT_SECTOR, _lib.VARIANT_RESIDENT):
        fs.operator.set_variant(variant)
        for prec in ("fp64", "mixed"):
            say("resident agreement", variant, prec)
            res[variant, prec] = (np.asarray(fs.solve_static(tol=1e-9, precision=prec)),
                                  dict(fs.last_solve_stats))
    for prec in ("fp64", "mixed"):
        (Us, ss), (Ur, sr) = res[_lib.VARIANT_SECTOR, prec], res[_lib.VARIANT_RESIDENT, prec]
        d = np.linalg.norm(Ur - Us) / np.linalg.norm(Us)
        say(prec, "iters", ss["iterations"], sr["iterations"], "launches", ss["launches"], sr["launches"], "diff", d)
        assert d < 1e-7
        if prec == "fp64":
            assert abs(sr["iterations"] - ss["iterations"]) <= 2
            assert sr["launches"] == 1
        else:
            assert sr["true_rel_residual"] < 1e-9


@pytest.mark.parametrize("variant", [5, 7])
def test_refined_fp32_and_fp64_element_products(h8, gpu, variant):
    """The refined solve with FP32 element products (arith 32, the library's
    choice) and with FP64 products of the FP32 data (arith 64): both meet the
    FP64 true residual and agree with the direct solve, and the two legs run
    different kernels (their results are not bitwise equal)."""
    say("refined arith", variant)
    rng = np.random.default_rng(11)
    nx, ny, nz = 24, 6, 6
    E = 10 ** rng.uniform(9, 12, nx * ny * nz)
    fs = _fesys(h8, nx, ny, nz, E_elem=E)
    fs.operator.set_variant(variant)
    c = o.cantilever(nx, ny, nz, E_elem=E)
    Ud = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    free, _ = o.partition(c["K"].shape[0], c["fixed_dofs"])
    its = {}
    sols = {}
    for arith in (32, 64):
        fs.operator.set_arith(arith)
        U = np.asarray(fs.solve_static(tol=1e-9, precision="mixed", preconditioner="jacobi"))
        sols[arith] = U
        st = fs.last_solve_stats
        true_res = np.linalg.norm((c["F"] - c["K"] @ U)[free]) / np.linalg.norm(c["F"][free])
        err = np.linalg.norm(U - Ud) / np.linalg.norm(Ud)
        its[arith] = (st["inner_iterations"], st["outer"])
        say("arith", arith, "fp32 iters", st["inner_iterations"], "replacements", st["outer"], "true_res", true_res,
            "err", err)
        assert true_res < 1e-9
        assert err < 1e-8
    fs.operator.set_arith(0)
    say("iterations by arithmetic", its)
    # a graph captured for one arithmetic is never replayed for the other
    assert not np.array_equal(sols[32], sols[64])


def test_nodal_and_heterogeneous_golden(h8, gpu, golden):
    fs = _fesys(h8, 20, 4, 4, load="nodal")
    U = fs.solve_static(method="pcg", tol=1e-8)
    assert abs(fs.last_solve_stats["iterations"] - int(golden["c1_nodal_1e-8/iters"])) <= 2
    E = golden["small_hetE_1e-8/E_elem"]
    fs = _fesys(h8, 8, 2, 2, E_elem=E)
    U = fs.solve_static(tol=1e-8)
    say("het iters", fs.last_solve_stats["iterations"], int(golden["small_hetE_1e-8/iters"]))
    assert abs(fs.last_solve_stats["iterations"] - int(golden["small_hetE_1e-8/iters"])) <= 3
    ref = golden["small_hetE_1e-8/U"]
    assert np.linalg.norm(np.asarray(U) - ref) / np.linalg.norm(ref) < 1e-7


def test_mixed_precision_refined_reaches_fp64_accuracy(h8, gpu):
    say("refined")
    rng = np.random.default_rng(2)
    nx, ny, nz = 24, 6, 6
    E = 10 ** rng.uniform(9, 12, nx * ny * nz)
    fs = _fesys(h8, nx, ny, nz, E_elem=E)
    U = np.asarray(fs.solve_static(tol=1e-8, precision="mixed"))
    st = fs.last_solve_stats
    c = o.cantilever(nx, ny, nz, E_elem=E)
    Ud = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    free, _ = o.partition(c["K"].shape[0], c["fixed_dofs"])
    true_res = np.linalg.norm((c["F"] - c["K"] @ U)[free]) / np.linalg.norm(c["F"][free])
    err = np.linalg.norm(U - Ud) / np.linalg.norm(Ud)
    say("refined stats", st, "true_res", true_res, "err", err)
    assert true_res < 1e-8
    assert abs(true_res - st["true_rel_residual"]) < 1e-3 * max(true_res, 1e-12) + 1e-13
    assert err < 1e-8
    U64 = np.asarray(fs.solve_static(tol=1e-8))
    say("fp64 iters", fs.last_solve_stats["iterations"])
    assert np.linalg.norm(U - U64) / np.linalg.norm(Ud) < 1e-8


def test_nonhomogeneous_dirichlet_and_reactions(h8, gpu):
    say("dirichlet")
    nx, ny, nz = 6, 3, 3
    mesh = h8.box_mesh(6.0, 3.0, 3.0, nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D())
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1000.0, nu=0.3)))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.fix_dofs(mesh.nodes_on_plane(0, 6.0), [0], value=0.01)
    U = np.asarray(fs.solve_static(method="pcg", tol=1e-12))
    c = o.cantilever(nx, ny, nz, E=1000.0)
    fixed = sorted(fs.fixed_dofs)
    vals = {d: fs.fixed_dof_values[d] for d in fixed}
    Ud = o.solve_direct(c["K"], np.zeros(c["K"].shape[0]), fixed, vals)
    assert np.linalg.norm(U - Ud) / np.linalg.norm(Ud) < 1e-9
    R = np.asarray(fs.reactions(U))
    Rref = c["K"] @ Ud
    assert np.abs(R[fixed] - Rref[fixed]).max() < 1e-8 * np.abs(Rref).max()
    assert abs(R[0::3].sum()) < 1e-8 * np.abs(R).max()       # equilibrium: x reactions balance


def test_error_paths(h8, gpu):
    say("errors")
    mesh = h8.box_mesh(4.0, 2.0, 2.0, 4, 2, 2)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D())
    with pytest.raises(RuntimeError):
        fs.solve_static()
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.add_nodal_force(mesh.nodes_on_plane(0, 4.0), 2, -1.0)
    with pytest.raises(RuntimeError, match="did not converge"):
        fs.solve_static(tol=1e-14, maxiter=3)
    with pytest.raises(ValueError, match="preconditioner"):
        fs.solve_static(method="pcg", preconditioner="bogus")
    # an indefinite operator (negative element scales) trips p^T A p <= 0
    fs2 = h8.FESystem(mesh, h8.Hex8Solid3D())
    fs2.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), element_scale=-np.ones(16))
    fs2.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs2.add_nodal_force(mesh.nodes_on_plane(0, 4.0), 2, -1.0)
    with pytest.raises(np.linalg.LinAlgError):
        fs2.solve_static(method="cg")
    # zero Jacobi diagonal at a free DOF (all elements around it have zero scale)
    fs3 = h8.FESystem(mesh, h8.Hex8Solid3D())
    fs3.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), element_scale=np.zeros(16))
    fs3.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs3.add_nodal_force(mesh.nodes_on_plane(0, 4.0), 2, -1.0)
    with pytest.raises(ValueError, match="zero diagonal"):
        fs3.solve_static()
    # zero load: warning and the trivial solution in 0 iterations
    fs4 = h8.FESystem(mesh, h8.Hex8Solid3D())
    fs4.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)))
    fs4.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    with pytest.warns(UserWarning):
        U = fs4.solve_static()
    assert np.all(np.asarray(U) == 0) and fs4.last_solve_stats["iterations"] == 0


def test_lazy_K_matches_oracle(h8, gpu):
    fs = _fesys(h8, 5, 3, 2)
    c = o.cantilever(5, 3, 2)
    assert abs(fs.K - c["K"]).max() < 1e-12 * abs(c["K"]).max()


def test_autotune_choice_cache_and_solution(h8, gpu, tmp_path, monkeypatch):
    say("autotune")
    import json
    from h8pcg import autotune
    monkeypatch.setenv("H8PCG_TUNE_CACHE", str(tmp_path / "tune.json"))
    nx, ny, nz = 24, 6, 6
    fs = h8.FESystem(h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz), h8.Hex8Solid3D())
    E = 10 ** np.random.default_rng(3).uniform(9, 11, nx * ny * nz)
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), element_scale=E)
    fs.fix_dofs(fs.mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.add_nodal_force(fs.mesh.nodes_on_plane(0, float(nx)), 2, -1e3)
    u_default = fs.solve_static_device(tol=1e-11).cpu().numpy()
    it_default = fs.last_solve_stats["iterations"]
    res = fs.operator.autotune()
    cands = set(autotune.CANDIDATES) | {(h8._lib.VARIANT_RESIDENT, 0)}
    for p in (32, 64):
        v, l, us = res[p]
        assert (v, l) in cands and us > 0
    data = json.loads((tmp_path / "tune.json").read_text())
    assert len(data) == 2
    assert fs.operator.autotune() == res            # served from the cache
    u = fs.solve_static_device(tol=1e-11).cpu().numpy()
    assert np.linalg.norm(u - u_default) < 1e-9 * np.linalg.norm(u_default)
    assert abs(fs.last_solve_stats["iterations"] - it_default) <= 2
    u32 = fs.solve_static_device(tol=1e-9, precision="mixed").cpu().numpy()
    assert np.linalg.norm(u32 - u_default) < 1e-7 * np.linalg.norm(u_default)


def test_refined_stagnation_is_reported(h8, gpu):
    """An unreachable tolerance ends the FP32-refined solve with a clear error
    once the true residual stops falling (not after maxiter FP32 iterations)."""
    say("refined stagnation")
    nx, ny, nz = 12, 4, 4
    fs = h8.FESystem(h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz), h8.Hex8Solid3D())
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=210e9, nu=0.3)))
    fs.fix_dofs(fs.mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.add_nodal_force(fs.mesh.nodes_on_plane(0, float(nx)), 2, -1e6)
    with pytest.raises(RuntimeError, match="stagnated"):
        fs.solve_static_device(tol=1e-18, precision="mixed", maxiter=10**7)


def test_free_body_minimum_norm_and_under_constrained(h8, gpu):
    """fe/tests/test_singular_system.py:61-162 and fe/tests/test_api_safety.py:
    214-236 on H8.  A free body (no supports) under equal and opposite end
    tension solves to the minimum-norm solution of the oracle's eigen_solve
    (K U = F, no mean drift, the bar elongates), in both precisions; with only
    u_x fixed on one face 3 rigid modes stay free and the result again equals
    eigen_solve on Kff; an unbalanced load raises LinAlgError; a prescribed
    displacement with zero load does not warn."""
    import warnings
    say("free body")
    nx, ny, nz, Lx = 6, 3, 3, 4.0
    mesh = h8.box_mesh(Lx, 1.0, 1.0, nx, ny, nz)
    nodes, elems = o.box_mesh(Lx, 1.0, 1.0, nx, ny, nz)
    K = o.assemble_csr(nodes, elems, o.d_solid3d(210e9, 0.3))
    left, right = mesh.nodes_on_plane(0, 0.0), mesh.nodes_on_plane(0, Lx)

    def system():
        fs = h8.FESystem(mesh, h8.Hex8Solid3D())
        fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=210e9, nu=0.3)))
        return fs

    fs = system()
    fs.add_nodal_force(right, 0, 1000.0)
    fs.add_nodal_force(left, 0, -1000.0)
    F = np.asarray(fs.F)
    Uref, nr = o.eigen_solve(K, F)
    assert nr == 6
    for prec in ("fp64", "mixed"):
        U = np.asarray(fs.solve_static(tol=1e-10, precision=prec))
        say("free body", prec, fs.last_solve_stats["iterations"])
        assert np.max(np.abs(U - Uref)) < 1e-8 * np.max(np.abs(Uref))
        assert np.max(np.abs(K @ U - F)) < 1e-6 * np.max(np.abs(F))
        for c in range(3):
            assert abs(np.mean(U[c::3])) < 1e-6 * np.max(np.abs(U))
        assert np.mean(U[3 * right]) - np.mean(U[3 * left]) > 0

    # partial support: u_x on the x = 0 face (translations y, z and the rotation
    # about x stay free); an end load in x is balanced against them
    fs2 = system()
    fs2.fix_dofs(left, [0])
    fs2.add_nodal_force(right, 0, 1000.0)
    U2 = np.asarray(fs2.solve_static(tol=1e-10))
    fixed = fs2.fixed_dofs_array
    free, fx = o.partition(K.shape[0], fixed)
    Kff, Ff = o.reduced_system(K, np.asarray(fs2.F), free, fx)
    uf, nr2 = o.eigen_solve(Kff, Ff)
    assert nr2 == 3
    assert np.max(np.abs(U2[free] - uf)) < 1e-8 * np.max(np.abs(uf))
    assert np.all(U2[fixed] == 0.0)

    # unbalanced loads on a free body
    fs.add_nodal_force(right, 1, -10.0)
    with pytest.raises(np.linalg.LinAlgError, match="under-constrained"):
        fs.solve_static()
    fs3 = system()
    fs3.add_nodal_force(right, 0, 1000.0)
    with pytest.raises(np.linalg.LinAlgError, match="(?i)not self-equilibrated"):
        fs3.solve_static()

    # prescribed displacement, zero load: no warning, non-zero field
    fs4 = system()
    fs4.fix_dofs(left, [0, 1, 2])
    fs4.fix_dofs(right, [0], value=1e-4)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        U4 = np.asarray(fs4.solve_static())
    assert np.any(U4)


def test_method_dispatch_matches_direct(h8, gpu):
    """fe/tests/test_iterative_method_dispatch.py:59-162 on an H8 cantilever
    (8x4x2, steel, tip shear): method None / 'pcg' / 'cg' agree with the
    oracle's direct solve; method=None is bit-identical to method='pcg' (the
    default); 'cg' honours a non-zero prescribed value exactly; the guards
    raise as the reference's do ('jacobi' is not a method)."""
    say("dispatch")
    nx, ny, nz = 8, 4, 2
    mesh = h8.box_mesh(2.0, 1.0, 0.5, nx, ny, nz)
    nodes, elems = o.box_mesh(2.0, 1.0, 0.5, nx, ny, nz)
    K = o.assemble_csr(nodes, elems, o.d_solid3d(2.1e11, 0.3))
    left, tip = mesh.nodes_on_plane(0, 0.0), mesh.nodes_on_plane(0, 2.0)

    def build(prescribed=False):
        fs = h8.FESystem(mesh, h8.Hex8Solid3D())
        fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=2.1e11, nu=0.3)))
        fs.fix_dofs(left, [0, 1, 2])
        if prescribed:
            fs.fix_dofs(tip, [0], value=0.02)
        else:
            fs.add_nodal_force(tip, 1, -1000.0)
        return fs

    fs = build()
    Uref = o.solve_direct(K, np.asarray(fs.F), fs.fixed_dofs_array)
    U_none = np.asarray(build().solve_static(tol=1e-10))
    U_pcg = np.asarray(build().solve_static(method="pcg", tol=1e-10))
    U_cg = np.asarray(build().solve_static(method="cg", tol=1e-10, maxiter=20000))
    assert np.array_equal(U_none, U_pcg)
    for U in (U_pcg, U_cg):
        assert np.max(np.abs(U - Uref)) < 1e-8 * np.max(np.abs(Uref))

    fs2 = build(prescribed=True)
    U2 = np.asarray(fs2.solve_static(method="cg", tol=1e-12))
    fixed = fs2.fixed_dofs_array
    U2ref = o.solve_direct(K, np.zeros(K.shape[0]), fixed, {int(d): fs2.fixed_dof_values[int(d)] for d in fixed})
    assert np.max(np.abs(U2 - U2ref)) < 1e-8 * np.max(np.abs(U2ref))
    assert np.all(U2[3 * tip] == 0.02)

    for bad in ("not_a_real_method", "jacobi"):
        with pytest.raises(ValueError, match="unknown method"):
            fs.solve_static(method=bad)
    with pytest.raises(ValueError, match="unknown preconditioner"):
        fs.solve_static(method="pcg", preconditioner="not_a_real_preconditioner")
    for pre in ("ssor", "ic0"):
        with pytest.raises(NotImplementedError):
            fs.solve_static(method="pcg", preconditioner=pre)
    # fe/tests/test_torch_sparse_solver.py:97-125: the torch-path drop-in on the
    # same cantilever agrees with the direct solve
    U_t = h8.fesystem_solve_static_torch(build(), method="cg", tol=1e-10)
    assert np.max(np.abs(U_t - Uref)) < 1e-6 * np.max(np.abs(Uref))
    assert np.all(U_t[fs.fixed_dofs_array] == 0.0)
    with pytest.raises(NotImplementedError):
        h8.fesystem_solve_static_torch(build(), method="dense")
    with pytest.raises(ValueError, match="unknown method"):
        h8.fesystem_solve_static_torch(build(), method="lu")
    with pytest.raises(NotImplementedError, match="prescribed"):
        h8.fesystem_solve_static_torch(build(prescribed=True))


def test_refined_warm_start(h8, gpu):
    """h8_refine_solve_x0 (the x0 of preconditioned_cg, fe/iterative_solvers.py:
    362-469, for the FP32-refined solve): from the FP64 solution the first true
    residual already meets tol, so no FP32 iteration runs and u is the guess;
    from a perturbed guess (garbage at the fixed DOFs, which are ignored) it
    converges in fewer FP32 iterations than from zero, to the same tolerance."""
    say("refined warm start")
    nx, ny, nz = 24, 6, 6
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D())
    E = 10 ** np.random.default_rng(3).uniform(9, 11, nx * ny * nz)
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), element_scale=E)
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.add_nodal_force(mesh.nodes_on_plane(0, float(nx)), 2, -1e6)
    U64 = np.asarray(fs.solve_static(tol=1e-12))
    Uc = np.asarray(fs.solve_static(precision="mixed", tol=1e-8))
    it_cold = fs.last_solve_stats["iterations"]
    Uw = np.asarray(fs.solve_static(precision="mixed", tol=1e-8, x0=U64))
    assert fs.last_solve_stats["iterations"] == 0 and np.array_equal(Uw, U64)
    rng = np.random.default_rng(4)
    x0 = U64 + 1e-3 * np.abs(U64).max() * rng.standard_normal(U64.shape)
    x0[fs.fixed_dofs_array] = 1.0
    Up = np.asarray(fs.solve_static(precision="mixed", tol=1e-8, x0=x0))
    st = fs.last_solve_stats
    say("cold", it_cold, "warm", st["iterations"])
    assert 0 < st["iterations"] < it_cold
    assert st["true_rel_residual"] <= 1e-8
    assert np.all(Up[fs.fixed_dofs_array] == 0.0)
    ref = np.abs(U64).max()
    assert np.abs(Up - U64).max() < 1e-5 * ref and np.abs(Uc - U64).max() < 1e-5 * ref


def test_resident_bricks_split_across_j_agree_with_streaming(h8, gpu):
    """On a 48x24x12 grid the resident planner splits j into two bricks (both
    precisions), so the z exchange crosses faces normal to j (scattered words,
    unlike bricks that span j).  The resident solve equals the streaming sector
    solve: FP64 to 1e-7 within two iterations, the refined form at its true
    residual."""
    from h8pcg import _lib
    rng = np.random.default_rng(11)
    nx, ny, nz = 48, 24, 12
    E = 10 ** rng.uniform(9, 11, nx * ny * nz)
    fs = _fesys(h8, nx, ny, nz, E_elem=E)
    res = {}
    for variant in (_lib.VARIANT_SECTOR, _lib.VARIANT_RESIDENT):
        fs.operator.set_variant(variant)
        for prec in ("fp64", "mixed"):
            say("split-j agreement", variant, prec)
            res[variant, prec] = (np.asarray(fs.solve_static(tol=1e-9, precision=prec)), dict(fs.last_solve_stats))
    for prec in ("fp64", "mixed"):
        (Us, ss), (Ur, sr) = res[_lib.VARIANT_SECTOR, prec], res[_lib.VARIANT_RESIDENT, prec]
        d = np.linalg.norm(Ur - Us) / np.linalg.norm(Us)
        say(prec, "iters", ss["iterations"], sr["iterations"], "diff", d)
        assert d < 1e-7
        if prec == "fp64":
            assert abs(sr["iterations"] - ss["iterations"]) <= 2 and sr["launches"] == 1
        else:
            assert sr["true_rel_residual"] < 1e-9


def test_diagonal_cache_follows_scales_and_supports(h8, gpu):
    """Repeated solves of one system reuse the Jacobi diagonal (no diagonal
    launch, no host wait); a new element scale (set_scale, update_scale_ or a
    direct s_e assignment) or a new support set rebuilds it -- in place, so
    captured graphs stay valid -- and each solve matches the oracle's direct
    solve of its own system."""
    import torch
    from h8pcg import _lib
    nx, ny, nz = 10, 3, 3
    rng = np.random.default_rng(11)
    E = 10 ** rng.uniform(9, 11, nx * ny * nz)
    fs = _fesys(h8, nx, ny, nz, E_elem=E)
    op = fs.operator
    c = o.cantilever(nx, ny, nz, E_elem=E)
    Ud = o.solve_direct(c["K"], c["F"], c["fixed_dofs"])
    L = _lib.lib()
    orig = L.h8_diag_f64
    calls = []

    def counted(*args):
        calls.append(1)
        return orig(*args)

    L.h8_diag_f64 = counted
    try:
        for prec in ("fp64", "mixed"):
            U1 = fs.solve_static_device(tol=1e-10, precision=prec).cpu().numpy()
            n0 = len(calls)
            U2 = fs.solve_static_device(tol=1e-10, precision=prec).cpu().numpy()
            assert len(calls) == n0, "an unchanged system rebuilt its diagonal"
            for U in (U1, U2):
                assert np.linalg.norm(U - Ud) / np.linalg.norm(Ud) < 1e-8
        ptr = op.dinv.data_ptr()
        d1 = op.dinv.clone()
        # scales doubled through set_scale: half the diagonal's inverse, half the displacement
        s0 = op.s_e.cpu().numpy()
        op.set_scale(2.0 * s0)
        U = fs.solve_static_device(tol=1e-10, precision="mixed").cpu().numpy()
        assert op.dinv.data_ptr() == ptr
        assert torch.equal(op.dinv, d1 / 2)
        assert np.linalg.norm(U - 0.5 * Ud) / np.linalg.norm(0.5 * Ud) < 1e-8
        # in place (the SIMP problem's way), then by direct assignment
        op.update_scale_(4.0 * s0)
        U = fs.solve_static_device(tol=1e-10, precision="fp64").cpu().numpy()
        assert torch.equal(op.dinv, d1 / 4)
        assert np.linalg.norm(U - 0.25 * Ud) / np.linalg.norm(0.25 * Ud) < 1e-8
        op.s_e = torch.as_tensor(8.0 * s0, device=op.device)
        op.s_e32 = op.s_e.float()
        U = fs.solve_static_device(tol=1e-10, precision="fp64").cpu().numpy()
        assert torch.equal(op.dinv, d1 / 8)
        assert np.linalg.norm(U - 0.125 * Ud) / np.linalg.norm(0.125 * Ud) < 1e-8
        # one more support set (the x = nx face's y components): a new diagonal
        op.set_scale(s0)
        tip = fs.mesh.nodes_on_plane(0, float(nx))
        fs.fix_dofs(tip, [1])
        n0 = len(calls)
        U = fs.solve_static_device(tol=1e-10, precision="mixed").cpu().numpy()
        assert len(calls) == n0 + 1
        tipy = [3 * int(n) + 1 for n in tip]
        assert float(op.dinv[tipy].abs().max()) == 0.0
        fixed = sorted(set(c["fixed_dofs"]) | set(tipy))
        Ut = o.solve_direct(c["K"], c["F"], fixed)
        assert np.linalg.norm(U - Ut) / np.linalg.norm(Ut) < 1e-8
    finally:
        L.h8_diag_f64 = orig
    say("diagonal cache ok")
