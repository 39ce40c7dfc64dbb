"""GPU parity of the SIMP path: device scale, closed-form and adjoint
sensitivities (autograd Functions) against the oracle and against
cancellation-free central differences of the forward solve."""
import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu


def case(nx=12, ny=4, nz=4, seed=3, p=3.0, rmin=1e-9):
    nodes, elems = o.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    ke = o.hex8_ke(nodes[elems[0]], o.d_solid3d(1.0, 0.3))
    rho = np.random.default_rng(seed).uniform(0.3, 1.0, len(elems))
    F = np.zeros(3 * len(nodes))
    tip = np.where((np.abs(nodes[:, 0] - nx) < 1e-9) & (np.abs(nodes[:, 1] - ny / 2) < 1e-9) & (nodes[:, 2] < 1e-9))[0]
    assert len(tip) == 1
    F[3 * tip + 2] = -1.0
    clamp = o.nodes_on_plane(nodes, 0, 0.0)
    fixed = [3 * int(n) + c for n in clamp for c in range(3)]
    s = o.simp_scale(rho, p, rmin)
    K = o.assemble_csr(nodes, elems, None, scale=s, ke=np.broadcast_to(ke, (len(elems), 24, 24)))
    u = o.solve_direct(K, F, fixed)
    return dict(nodes=nodes, elems=elems, ke=ke, rho=rho, F=F, fixed=fixed, K=K, u=u, p=p, rmin=rmin, dims=(nx, ny, nz))


def problem(c, gpu, **kw):
    from h8pcg.topopt import SIMPProblem
    return SIMPProblem(*c["dims"], c["F"], c["fixed"], p=c["p"], rho_min=c["rmin"], device=gpu, **kw)


def test_simp_scale_kernel(gpu):
    import torch
    from h8pcg.topopt import simp_scale
    rho = np.random.default_rng(0).uniform(0, 1, 1000)
    s, ds = simp_scale(torch.as_tensor(rho, device=gpu), 3.0, 1e-3, with_grad=True)
    assert np.allclose(s.cpu().numpy(), o.simp_scale(rho, 3.0, 1e-3), rtol=1e-14, atol=0)
    assert np.allclose(ds.cpu().numpy(), o.simp_scale_grad(rho, 3.0, 1e-3), rtol=1e-14, atol=0)


def test_compliance_and_closed_form_gradient(gpu):
    import torch
    from h8pcg import topopt
    c = case()
    pb = problem(c, gpu, tol=1e-12)
    rho = torch.as_tensor(c["rho"], device=gpu).requires_grad_(True)
    C = topopt.compliance(rho, pb)
    C.backward()
    g = rho.grad.cpu().numpy()
    C_ref = o.compliance(c["F"], c["u"])
    g_ref = o.compliance_sensitivity_closed_form(c["ke"], c["elems"], c["u"], c["rho"], c["p"], c["rmin"])
    print("[gpu-test] C", float(C), C_ref, "grad err", np.abs(g - g_ref).max() / np.abs(g_ref).max(), flush=True)
    assert abs(float(C) - C_ref) < 1e-9 * abs(C_ref)
    assert np.abs(g - g_ref).max() < 1e-8 * np.abs(g_ref).max()
    # recomputing the forward instead of storing it gives the same gradient:
    # forced, and by the size rule (a working-set cap below what is allocated)
    assert pb.keep_forward(torch.zeros(10, dtype=torch.float64, device=gpu))
    for kw in (dict(store_forward=False), dict(working_set_bytes=1)):
        pb2 = problem(c, gpu, tol=1e-12, **kw)
        assert not pb2.keep_forward(torch.zeros(10, dtype=torch.float64, device=gpu))
        rho2 = torch.as_tensor(c["rho"], device=gpu).requires_grad_(True)
        solves = pb2.solves
        topopt.compliance(rho2, pb2).backward()
        assert pb2.solves == solves + 2          # the forward, and its recomputation in backward
        assert np.abs(rho2.grad.cpu().numpy() - g).max() < 1e-9 * np.abs(g).max()
    with pytest.raises(ValueError):
        problem(c, gpu, store_forward="sometimes")


def test_general_adjoint_displacement_objective(gpu):
    import torch
    from h8pcg import topopt
    c = case(10, 4, 4, seed=5)
    pb = problem(c, gpu, tol=1e-12)
    rng = np.random.default_rng(1)
    w = rng.normal(size=len(c["F"]))
    rho = torch.as_tensor(c["rho"], device=gpu).requires_grad_(True)
    u = topopt.displacement(rho, pb)
    L = (torch.as_tensor(w, device=gpu) * u).sum()
    L.backward()
    g = rho.grad.cpu().numpy()
    # oracle: K lam = -dL/du, dL/drho_e = s' lam_e^T ke u_e
    g_ref = o.adjoint_sensitivity(c["K"], c["ke"], c["elems"], c["fixed"], c["u"], w, c["rho"], c["p"], c["rmin"])
    print("[gpu-test] adjoint err", np.abs(g - g_ref).max() / np.abs(g_ref).max(), flush=True)
    assert np.abs(g - g_ref).max() < 1e-7 * np.abs(g_ref).max()
    # compliance through the general path equals the closed form
    rho3 = torch.as_tensor(c["rho"], device=gpu).requires_grad_(True)
    u3 = topopt.displacement(rho3, pb)
    (torch.as_tensor(c["F"], device=gpu) * u3).sum().backward()
    g_cf = o.compliance_sensitivity_closed_form(c["ke"], c["elems"], c["u"], c["rho"], c["p"], c["rmin"])
    assert np.abs(rho3.grad.cpu().numpy() - g_cf).max() < 1e-7 * np.abs(g_cf).max()


def test_gradient_matches_central_differences(gpu):
    import torch
    from h8pcg import topopt
    c = case(16, 4, 4, seed=7)
    pb = problem(c, gpu, tol=1e-13)
    rho = torch.as_tensor(c["rho"], device=gpu).requires_grad_(True)
    topopt.compliance(rho, pb).backward()
    g = rho.grad.cpu().numpy()
    els = np.random.default_rng(0).choice(len(c["rho"]), 12, replace=False)
    fd, _ = topopt.fd_check(pb, rho.detach(), els, h=1e-6, tol=1e-13)
    rel = np.abs(fd - g[els]) / np.abs(g[els])
    print("[gpu-test] fd rel max", rel.max(), flush=True)
    assert rel.max() < 1e-6


def test_mixed_forward_gradient(gpu):
    import torch
    from h8pcg import topopt
    c = case(12, 4, 4, seed=9)
    pb = problem(c, gpu, tol=1e-10, precision="mixed")
    rho = torch.as_tensor(c["rho"], device=gpu).requires_grad_(True)
    topopt.compliance(rho, pb).backward()
    g_ref = o.compliance_sensitivity_closed_form(c["ke"], c["elems"], c["u"], c["rho"], c["p"], c["rmin"])
    assert np.abs(rho.grad.cpu().numpy() - g_ref).max() < 1e-7 * np.abs(g_ref).max()


@pytest.mark.parametrize("dims", [(3, 2, 5), (37, 21, 19)])
def test_element_bilinear_matches_oracle(gpu, dims):
    """h8_elem_bilinear_f64 (the marching sector form) against the oracle's
    per-element u_e^T K0 v_e, with partial column tiles and several slab
    chunks; u == v (one staged vector) and u != v."""
    import torch
    from h8pcg.operator import GridOperator
    from h8pcg.elements import Hex8Solid3D
    nx, ny, nz = dims
    nodes, elems = o.box_mesh(nx * 1.0, ny * 0.5, nz * 0.75, nx, ny, nz)
    X = nodes[elems[0]]
    K0 = Hex8Solid3D().stiffness_batched(X[None], o.d_solid3d(1.0, 0.3), device=gpu)[0]
    m = len(elems)
    op = GridOperator(nx, ny, nz, K0, np.ones(m), device=gpu)
    ke = op.K0.cpu().numpy()
    rng = np.random.default_rng(sum(dims))
    u = rng.standard_normal(3 * len(nodes))
    v = rng.standard_normal(3 * len(nodes))
    w = rng.standard_normal(m)
    dofs = (3 * elems[:, :, None] + np.arange(3)).reshape(m, 24)
    ut = torch.as_tensor(u, device=gpu)
    for a, b, bt in ((u, v, torch.as_tensor(v, device=gpu)), (u, u, ut), (u, u, ut.clone())):
        ref = w * np.einsum("ei,ij,ej->e", a[dofs], ke, b[dofs])
        got = op.element_bilinear(ut, bt, w).cpu().numpy()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print("[gpu-test] bilinear", dims, "same" if bt is ut else "two", err, flush=True)
        assert err < 1e-13
