"""GPU parity of the batched solves: every (instance, load case) system against
the oracle's direct solve, per-system convergence masks, and the multi-RHS
FESystem mirror of solve_static_batched."""
import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu


FORMS = [0, 1, 2]   # launched (graphs of two kernels per iteration), persistent 8x16, persistent 16x16


@pytest.mark.parametrize("form", FORMS)
def test_solve_static_batched_matches_oracle(gpu, form):
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
    U, st = solve_static_batched(fs, Fb, tol=1e-11, return_stats=True, form=form)
    print("[gpu-test] batched stats", st["form"], st["iterations"], st["elapsed_ms"], flush=True)
    c = o.cantilever(nx, ny, nz, E_elem=E, load="nodal", total=0.0)
    fixed = sorted(fs.fixed_dofs)
    vals = {d: fs.fixed_dof_values[d] for d in fixed}
    Uref = o.solve_static_batched(c["K"], Fb, fixed, vals)
    for j in range(5):
        err = np.linalg.norm(U[:, j] - Uref[:, j]) / np.linalg.norm(Uref[:, j])
        assert err < 1e-8, (j, err)
    assert np.all(st["status"] == 0)


@pytest.mark.parametrize("form", FORMS)
def test_uneven_instances_and_masks(gpu, form):
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
    X, st = solve_many(systems, tol=1e-10, form=form)
    print("[gpu-test] solve_many", st["form"], st["iterations"].tolist(), st["elapsed_ms"], st["launches"], flush=True)
    if form != 0:
        assert st["launches"] == 1
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


def test_batched_dirichlet_reference_checks(gpu):
    """fe/tests/test_batched_nonhomogeneous_dirichlet.py:53-147 on H8: the
    prescribed value appears in every column; a one-column batch equals
    solve_static; an explicit value=0.0 gives bit-identical results to an
    omitted value."""
    import h8pcg
    from h8pcg.batched import solve_static_batched
    nx, ny, nz = 8, 3, 2
    mesh = h8pcg.box_mesh(2.0, 1.0, 0.5, nx, ny, nz)
    left, right = mesh.nodes_on_plane(0, 0.0), mesh.nodes_on_plane(0, 2.0)

    def build(value=None, zero_explicit=False):
        fs = h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D())
        fs.assemble_stiffness(h8pcg.D_solid3d(h8pcg.Material(E=2.1e11, nu=0.3)))
        if zero_explicit:
            fs.fix_dofs(left, [0, 1, 2], value=0.0)
        else:
            fs.fix_dofs(left, [0, 1, 2])
        if value is not None:
            fs.fix_dofs(right, [0], value=value)
        return fs

    rng = np.random.default_rng(0)
    fs = build(0.015)
    Fb = rng.normal(size=(fs.n_dof, 4)) * 1000.0
    U = solve_static_batched(fs, Fb, tol=1e-12)
    assert np.all(U[3 * right, :] == 0.015)

    F = np.zeros(fs.n_dof)
    F[3 * right + 1] = -500.0 / len(right)
    fs1 = build(0.01)
    U1 = solve_static_batched(fs1, F[:, None], tol=1e-10)
    fs_ref = build(0.01)
    fs_ref.F[:] = F
    Uref = np.asarray(fs_ref.solve_static(tol=1e-10))
    assert np.max(np.abs(U1[:, 0] - Uref)) < 1e-8 * np.max(np.abs(Uref))

    Fz = rng.normal(size=(fs.n_dof, 3)) * 1000.0
    Ua = solve_static_batched(build(), Fz)
    Ub = solve_static_batched(build(zero_explicit=True), Fz)
    assert np.array_equal(Ua, Ub)


def test_solve_many_zero_diagonal_and_columns(gpu):
    """solve_many checks every model's Jacobi diagonal once, after all of them
    (the jacobi_preconditioner ValueError of fe/iterative_solvers.py:193-202),
    and builds each model's right-hand sides in one masking launch: a model
    with a load case of zeros next to real ones solves each column as a single
    solve_static would."""
    import h8pcg
    from h8pcg.batched import solve_many
    mesh = h8pcg.box_mesh(4.0, 2.0, 2.0, 4, 2, 2)

    def model(scale):
        fs = h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D())
        fs.assemble_stiffness(h8pcg.D_solid3d(h8pcg.Material(E=1.0, nu=0.3)), element_scale=scale)
        fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
        return fs

    good, bad = model(np.ones(16)), model(np.zeros(16))
    F = np.zeros((good.n_dof, 3))
    tip = mesh.nodes_on_plane(0, 4.0)
    F[3 * tip + 2, 0] = -1.0
    F[3 * tip + 1, 2] = 0.5
    with pytest.raises(ValueError, match="zero diagonal"):
        solve_many([(good, F), (bad, F)], tol=1e-10)
    X, st = solve_many([(good, F)], tol=1e-10)
    U = X[0].cpu().numpy()
    assert np.all(U[:, 1] == 0.0) and np.all(st["status"] == 0)
    for c in (0, 2):
        ref = model(np.ones(16))
        ref.F[:] = F[:, c]
        Uc = np.asarray(ref.solve_static(tol=1e-10))
        assert np.max(np.abs(U[:, c] - Uc)) < 1e-8 * np.max(np.abs(Uc))


def test_batched_graph_replays_follow_tol_and_maxiter(gpu):
    """Two batched solves of the same shapes (so the same cached graph and, with
    torch's caching allocator, usually the same buffers) at tol 1e-6 and then
    1e-12: the second solve stops at its own tolerance, and a third call with a
    smaller iteration limit stops at that limit."""
    import h8pcg
    from h8pcg.batched import solve_many
    from h8pcg.workloads import cantilever, tip_traction_loads
    rng = np.random.default_rng(3)
    systems = []
    for nx, k in [(8, 3), (12, 2)]:
        E = 10 ** rng.uniform(9, 11, nx * (nx // 4) ** 2)
        fs = cantilever(nx, nx // 4, nx // 4, E_elem=E, load="none")
        systems.append((fs, tip_traction_loads(fs, nx, nx // 4, nx // 4, k, rng)))
    _, st6 = solve_many(systems, tol=1e-6)
    _, st12 = solve_many(systems, tol=1e-12)
    print("[gpu-test] rel 1e-6", st6["rel_residual"].max(), "rel 1e-12", st12["rel_residual"].max(), flush=True)
    assert np.all(st6["status"] == 0) and np.all(st12["status"] == 0)
    assert np.all(st6["rel_residual"] < 1e-6)
    assert np.all(st12["rel_residual"] < 1e-12)
    assert np.all(st12["iterations"] > st6["iterations"])
    with pytest.raises(RuntimeError, match="did not converge"):
        solve_many(systems, tol=1e-12, maxiter=5)


def test_batched_forms_agree_and_the_tuner_caches_its_choice(gpu, tmp_path, monkeypatch):
    """The three forms solve the same systems to the same tolerance with the
    same recurrence (iteration counts within 2: only the summation order of the
    dot products differs) and agree to 1e-9; the batched autotuner picks one of
    them, records it for this batch composition and serves it from the cache."""
    import json
    from h8pcg import batched
    from h8pcg.workloads import cantilever, tip_traction_loads
    monkeypatch.setenv("H8PCG_TUNE_CACHE", str(tmp_path / "tune.json"))
    rng = np.random.default_rng(11)
    systems = []
    for nx, k in [(24, 3), (8, 2), (16, 4), (12, 1)]:
        ny = nz = nx // 4
        E = 10 ** rng.uniform(9, 11, nx * ny * nz)
        fs = cantilever(nx, ny, nz, E_elem=E, load="none")
        systems.append((fs, tip_traction_loads(fs, nx, ny, nz, k, rng)))
    out = {}
    for form in FORMS:
        X, st = batched.solve_many(systems, tol=1e-10, form=form)
        assert np.all(st["status"] == 0)
        out[form] = ([x.cpu().numpy() for x in X], st["iterations"].copy())
    ref_x, ref_it = out[0]
    for form in FORMS[1:]:
        xs, its = out[form]
        assert np.max(np.abs(its - ref_it)) <= 2, (form, its, ref_it)
        for a, b in zip(xs, ref_x):
            assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b)
    X, st = batched.solve_many(systems, tol=1e-10)        # tuned
    data = json.load(open(tmp_path / "tune.json"))
    keys = [k for k in data if "|batched|" in k]
    print("[gpu-test] batched tuner", st["form"], data[keys[0]], flush=True)
    assert len(keys) == 1 and data[keys[0]][0] in FORMS
    assert st["form"] == batched.FORM_NAMES[data[keys[0]][0]]
    X2, st2 = batched.solve_many(systems, tol=1e-10)      # served from the cache
    assert st2["form"] == st["form"]
    assert all(np.array_equal(a.cpu().numpy(), b.cpu().numpy()) for a, b in zip(X, X2))


@pytest.mark.parametrize("form", [1, 2])
def test_persistent_form_stops_at_maxiter(gpu, form):
    from h8pcg import batched
    from h8pcg.workloads import cantilever, tip_traction_loads
    rng = np.random.default_rng(12)
    fs = cantilever(16, 4, 4, E_elem=10 ** rng.uniform(9, 11, 256), load="none")
    F = tip_traction_loads(fs, 16, 4, 4, 3, rng)
    with pytest.raises(RuntimeError):
        batched.solve_many([(fs, F)], tol=1e-12, maxiter=7, form=form)
    fs._op.set_free_mask(fs._free_mask())
    cols, _ = batched._dirichlet_columns(fs, F)
    b = batched._Batch(fs._op.device)
    b.add(fs._op, cols)
    _, st = b._solve(1e-12, 7, form, check_rc=False)
    assert np.all(st["iterations"] == 7) and np.all(st["status"] != 0)
