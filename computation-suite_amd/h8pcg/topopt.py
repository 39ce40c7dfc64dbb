"""SIMP compliance and its adjoint gradient on the GPU (fe/topopt.py:74-229,
fe/differentiable.py:478-612).

``SIMPProblem`` holds one structured H8 grid with a reference element K0 (E0,
nu), a load F and clamped DOFs.  K(rho) = sum_e s(rho_e) K0 with
s = rho_min + (1-rho_min) rho^p (simp_scale); every solve, scale and
element contraction runs in libh8pcg.so.

Two torch.autograd.Functions sit on top:
  * ``compliance(rho, problem)``  -> C = F.u.  Backward uses the self-adjoint
    closed form dC/drho_e = -s'(rho_e) u_e^T K0 u_e (no second solve), with u
    stored by the forward pass or recomputed by a second solve in backward.
    store_forward=True / False fixes the choice; "auto" (the default) stores
    u while the device's allocated bytes plus u stay within working_set_bytes
    (80 GiB by default) and recomputes it beyond that, so many forward passes
    awaiting their backward cannot push the working set past the cap.
  * ``displacement(rho, problem)`` -> u (full length).  Backward solves the
    adjoint K lambda = dL/du (masked) with the same PCG and returns
    dL/drho_e = -s'(rho_e) lambda_e^T K0 u_e.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import check, ptr
from .elements import Hex8Solid3D
from .material import D_solid3d, Material
from .operator import GridOperator, require_gpu


def simp_scale(rho, p=3.0, rho_min=1e-3, with_grad=False):
    """s(rho) (and s'(rho)) as device tensors (fe/topopt.py:74-90)."""
    import torch
    r = rho.detach().to(torch.float64).contiguous()
    s = torch.empty_like(r)
    ds = torch.empty_like(r) if with_grad else None
    with torch.cuda.device(r.device):
        check(_lib.lib().h8_simp_scale_f64(ptr(r), float(p), float(rho_min), r.numel(), ptr(s), ptr(ds),
                                           _lib.stream_handle(r.device)), "h8_simp_scale_f64")
    return (s, ds) if with_grad else s


class SIMPProblem:
    """Clamped structured H8 body under a fixed load, stiffness scaled per element."""

    def __init__(self, nx, ny, nz, F, fixed_dofs, E0=1.0, nu=0.3, h=(1.0, 1.0, 1.0), p=3.0, rho_min=1e-9,
                 device="cuda", tol=1e-10, precision="fp64", maxiter=None, store_forward="auto", D=None,
                 preconditioner=None, working_set_bytes=80 * 2**30):
        torch = require_gpu()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.m = self.nx * self.ny * self.nz
        self.p, self.rho_min = float(p), float(rho_min)
        self.tol, self.precision, self.maxiter = tol, precision, maxiter
        if store_forward not in (True, False, "auto"):
            raise ValueError("store_forward must be True, False or 'auto'")
        self.store_forward = store_forward
        self.working_set_bytes = int(working_set_bytes)
        # FP32 iterations (precision "mixed"): the multigrid V-cycle where the grid
        # coarsens far enough (None = that choice), else "jacobi"
        self.preconditioner = preconditioner
        X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                      [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float) * np.asarray(h, dtype=float)
        D0 = D_solid3d(Material(E=E0, nu=nu)) if D is None else np.asarray(D, dtype=float)
        K0 = Hex8Solid3D().stiffness_batched(X[None], D0, device=device)[0]
        self.op = GridOperator(nx, ny, nz, K0, np.ones(self.m), device=device)
        n = self.op.n_dof
        mask = np.ones(n, dtype=np.uint8)
        mask[np.asarray(list(fixed_dofs), dtype=np.int64)] = 0
        self.free_mask = mask
        self.F = torch.as_tensor(np.asarray(F, dtype=float), device=self.op.device)
        self.last_stats = None
        self.solves = 0

    @property
    def device(self):
        return self.op.device

    def keep_forward(self, u):
        """Whether a forward pass stores u for its backward (else recomputes it)."""
        if self.store_forward != "auto":
            return bool(self.store_forward)
        import torch
        need = u.numel() * u.element_size()
        return torch.cuda.memory_allocated(u.device) + need <= self.working_set_bytes

    def _set_rho(self, rho):
        s, ds = simp_scale(rho, self.p, self.rho_min, with_grad=True)
        self.op.update_scale_(s)
        self.op.set_free_mask(self.free_mask)
        return s, ds

    def solve(self, rho, rhs=None, x0=None, tol=None, precision=None):
        """u with K(rho) u = rhs (default F) on the free DOFs (precision: the
        problem's unless given; x0: warm start, in either precision)."""
        self._set_rho(rho)
        b = self.op.masked(self.F if rhs is None else rhs.to(torch_f64()).contiguous())
        tol = self.tol if tol is None else tol
        precision = self.precision if precision is None else precision
        if precision == "mixed":
            pc = self.preconditioner or ("mg" if self.op.mg_available() else "jacobi")
            u, st = self.op.refined(b, tol=tol, maxiter=self.maxiter, x0=x0, preconditioner=pc)
        else:
            u, st = self.op.pcg(b, tol=tol, maxiter=self.maxiter, x0=x0)
        self.last_stats = st
        self.solves += 1
        return u

    def compliance_value(self, u):
        return float((self.F * u).sum())

    def sensitivity_closed_form(self, rho, u):
        """dC/drho_e = -s'(rho_e) u_e^T K0 u_e (fe/topopt.py:161-177)."""
        _, ds = simp_scale(rho, self.p, self.rho_min, with_grad=True)
        return self.op.element_bilinear(u, u, -ds)

    def adjoint_vjp(self, rho, u, lam):
        """dL/drho_e = -s'(rho_e) lambda_e^T K0 u_e for K lambda = dL/du."""
        _, ds = simp_scale(rho, self.p, self.rho_min, with_grad=True)
        return self.op.element_bilinear(lam, u, -ds)


def sensitivity_filter(nx, ny, nz, h, radius, rho, dc, eps=1e-9):
    """Sigmund's filter of dc over the grid's elements (fe/topopt.py:235-252) on
    the device: sum_j H_ej rho_j dc_j / max(max(rho_e, eps) sum_j H_ej, eps) with
    H_ej = max(0, radius - |c_e - c_j|); h = brick spacing."""
    import torch
    r = rho.detach().to(torch.float64).contiguous()
    d = dc.detach().to(device=r.device, dtype=torch.float64).contiguous()
    if r.numel() != nx * ny * nz or d.numel() != r.numel():
        raise ValueError(f"rho and dc need {nx * ny * nz} entries")
    out = torch.empty_like(r)
    g = _lib.H8Grid(int(nx), int(ny), int(nz), 0, 0)
    with torch.cuda.device(r.device):
        check(_lib.lib().h8_sens_filter_f64(__import__("ctypes").byref(g), float(h[0]), float(h[1]), float(h[2]),
                                            float(radius), ptr(r), ptr(d), float(eps), ptr(out),
                                            _lib.stream_handle(r.device)), "h8_sens_filter_f64")
    return out


def oc_update(rho, dc, volume_fraction, move=0.2, rho_min_bound=1e-3, rho_max_bound=1.0, bisection_tol=1e-4,
              max_bisection_iter=100, return_change=False):
    """Optimality-criteria density update (fe/topopt.py:258-296) on the device:
    the reference's bracket widening and bisection on the volume multiplier,
    every pass one kernel over the elements.  Returns rho_new (device tensor),
    and max |rho_new - rho| when return_change."""
    import ctypes
    import torch
    r = rho.detach().to(torch.float64).contiguous()
    d = dc.detach().to(device=r.device, dtype=torch.float64).contiguous()
    m = r.numel()
    if d.numel() != m:
        raise ValueError("rho and dc must have the same length")
    out = torch.empty_like(r)
    ws = torch.empty(int(_lib.lib().h8_oc_workspace_bytes(m)), dtype=torch.uint8, device=r.device)
    change = ctypes.c_double(0.0)
    with torch.cuda.device(r.device):
        check(_lib.lib().h8_oc_update_f64(ptr(r), ptr(d), m, float(volume_fraction), float(move),
                                          float(rho_min_bound), float(rho_max_bound), float(bisection_tol),
                                          int(max_bisection_iter), ptr(out), ctypes.byref(change), ptr(ws),
                                          ws.numel(), _lib.stream_handle(r.device)), "h8_oc_update_f64")
    return (out, change.value) if return_change else out


def topology_optimize_compliance(fesystem, D, F_ext, volume_fraction, p=3.0, rho_min=1e-3, move=0.2, n_iter=50,
                                 filter_radius=None, sensitivity="closed_form", rho_init=None, tol=1e-3,
                                 solve_tol=1e-10, precision="fp64", timings=None):
    """SIMP compliance minimisation on the GPU (fe/topopt.py:302-377): solve ->
    sensitivity -> (filter) -> OC update, until max |rho_new - rho| < tol.
    fesystem supplies the structured mesh and the fixed DOFs (its own K is not
    used), D the rho = 1 material.  Every step runs in libh8pcg.so; the forward
    solve is the Jacobi-PCG to solve_tol (FP64, or precision="mixed"), warm-started
    from the previous iteration's u.  sensitivity="adjoint" takes the gradient
    through the general adjoint (a second solve) instead of the closed form.
    Returns {"rho", "compliance_history", "u_final"} as the reference does.
    timings: a dict that receives the device-synchronised seconds spent in each
    stage ("solve", "sensitivity", "filter", "oc"), summed over the iterations,
    and each iteration's seconds under timings["per_iteration"][stage] (None:
    no synchronisation between the stages)."""
    import time
    import torch
    clock = [time.perf_counter()]

    def lap(stage):
        if timings is not None:
            torch.cuda.synchronize()
            t = time.perf_counter()
            timings[stage] = timings.get(stage, 0.0) + t - clock[0]
            timings.setdefault("per_iteration", {}).setdefault(stage, []).append(t - clock[0])
            clock[0] = t
    from .mesh import structured_grid
    if sensitivity not in ("closed_form", "adjoint"):
        raise ValueError(f"topology_optimize_compliance: unknown sensitivity={sensitivity!r} "
                         f"-- expected 'closed_form' or 'adjoint'.")
    st = structured_grid(fesystem.mesh)
    if st is None:
        raise NotImplementedError("topology_optimize_compliance(backend='hip') needs a structured box mesh")
    nx, ny, nz, coords0 = st
    h = coords0[6] - coords0[0]
    prob = SIMPProblem(nx, ny, nz, F_ext, fesystem.fixed_dofs_array, h=h, p=p, rho_min=rho_min,
                       device=fesystem.device, tol=solve_tol, precision=precision, D=D)
    m = nx * ny * nz
    rho = (torch.full((m,), float(volume_fraction), dtype=torch.float64, device=prob.device) if rho_init is None
           else torch.as_tensor(np.asarray(rho_init, dtype=float), device=prob.device))
    history, u = [], None
    lap("setup")
    for _ in range(int(n_iter)):
        u = prob.solve(rho, x0=u)
        history.append(prob.compliance_value(u))
        lap("solve")
        if sensitivity == "closed_form":
            dc = prob.sensitivity_closed_form(rho, u)
        else:
            r = rho.clone().requires_grad_(True)
            with torch.enable_grad():
                (prob.F * displacement(r, prob)).sum().backward()
            dc = r.grad.detach()
        lap("sensitivity")
        if filter_radius is not None and filter_radius > 0:
            dc = sensitivity_filter(nx, ny, nz, h, filter_radius, rho, dc)
            lap("filter")
        rho_new, change = oc_update(rho, dc, volume_fraction, move=move, rho_min_bound=rho_min, return_change=True)
        lap("oc")
        rho = rho_new
        if change < tol:
            break
    return {"rho": rho.cpu().numpy(), "compliance_history": history, "u_final": u.cpu().numpy()}


def torch_f64():
    import torch
    return torch.float64


def _functions():
    import torch

    class _Compliance(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho, problem):
            u = problem.solve(rho)
            ctx.problem = problem
            ctx.stored = problem.keep_forward(u)
            if ctx.stored:
                ctx.save_for_backward(rho, u)
            else:
                ctx.save_for_backward(rho)
            return (problem.F * u).sum().reshape(())

        @staticmethod
        def backward(ctx, gC):
            problem = ctx.problem
            if ctx.stored:
                rho, u = ctx.saved_tensors
            else:
                (rho,) = ctx.saved_tensors
                u = problem.solve(rho)
            g = problem.sensitivity_closed_form(rho, u)
            return (gC * g).to(rho.dtype), None

    class _Displacement(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho, problem):
            u = problem.solve(rho)
            ctx.problem = problem
            ctx.save_for_backward(rho, u)
            return u

        @staticmethod
        def backward(ctx, gu):
            rho, u = ctx.saved_tensors
            problem = ctx.problem
            lam = problem.solve(rho, rhs=gu.contiguous())
            return problem.adjoint_vjp(rho, u, lam).to(rho.dtype), None

    return _Compliance, _Displacement


_FNS = None


def compliance(rho, problem):
    """C(rho) = F.u(rho) as a differentiable torch scalar."""
    global _FNS
    if _FNS is None:
        _FNS = _functions()
    return _FNS[0].apply(rho, problem)


def displacement(rho, problem):
    """u(rho) (full length) as a differentiable torch tensor (general adjoint backward)."""
    global _FNS
    if _FNS is None:
        _FNS = _functions()
    return _FNS[1].apply(rho, problem)


def fd_check(problem, rho, elements, h=1e-6, tol=1e-12):
    """Cancellation-free central differences of C at the given elements:
    C(rho+h e) - C(rho-h e) = -(s+ - s-) u-_e^T K0 u+_e (K symmetric), all three
    solves in FP64 to `tol`, the perturbed ones warm-started from u(rho).
    Returns (fd, u) as numpy."""
    import torch
    u = problem.solve(rho, tol=tol, precision="fp64")
    out = []
    s_all = simp_scale(rho, problem.p, problem.rho_min)
    for e in elements:
        rp = rho.clone(); rm = rho.clone()
        rp[e] += h
        rm[e] -= h
        up = problem.solve(rp, x0=u, tol=tol, precision="fp64")
        um = problem.solve(rm, x0=u, tol=tol, precision="fp64")
        sp_ = simp_scale(rp[e:e + 1], problem.p, problem.rho_min)
        sm_ = simp_scale(rm[e:e + 1], problem.p, problem.rho_min)
        w = torch.zeros(problem.m, dtype=torch.float64, device=problem.device)
        w[e] = -(sp_ - sm_)[0]
        dC = problem.op.element_bilinear(um, up, w)[e]
        out.append(float(dC) / (2 * h))
    problem._set_rho(rho)
    del s_all
    return np.array(out), u
