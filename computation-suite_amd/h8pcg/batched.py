"""Batched solves on the GPU.

``solve_static_batched(fs, F_batch)`` mirrors fe/batched_solve.py:105-164: one
assembled FESystem, many load vectors (columns of F_batch), the same
Dirichlet data (non-zero prescribed values corrected per column), a full
(n_dof, n_batch) result with U[fixed] = u_fixed.  Instead of one factorisation
it runs one batched Jacobi-PCG over all columns.

``solve_many(systems)`` solves many *different* structured models, each with
its own load cases, in the same batched launches (the role fe/batch.py:60-128
plays with a process pool): every (model, load case) pair is a system with
its own convergence mask.

Forms (h8_pcg_batched_form_f64): every system advanced together, two launches
per iteration in captured graphs ("launched"), or one persistent launch whose
workgroups take whole systems from a queue and run each to the end, with the
direction kernel's sector march on 8 x 16 or 16 x 16 element columns.  The
batched autotuner (``choose_form``) times the three on the batch itself (40
iterations of every system, tol 0) the first time a batch of that composition
is seen on a device, and keeps the fastest; the choice is cached with the
kernel autotuner's (h8pcg.autotune.cache_path()).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import H8Instance, check, ptr


def _dirichlet_columns(fs, F_batch):
    """Device b columns mask(F - K u_D) and the u_D vector (None if zero)."""
    import torch
    op = fs._op
    fixed = fs.fixed_dofs_array
    u_fixed = fs._fixed_dof_values_array(fixed)
    # columns as rows of one (k, n_dof) array: one upload and one masking launch
    F_t = torch.as_tensor(np.ascontiguousarray(np.asarray(F_batch, dtype=float).T), device=op.device)
    corr = None
    uD_t = None
    if len(fixed) and np.any(u_fixed):
        uD = np.zeros(fs.n_dof)
        uD[fixed] = u_fixed
        uD_t = torch.as_tensor(uD, device=op.device)
        corr = op.matvec(uD_t)
    B = op.masked_cols(F_t, corr)
    return list(B.unbind(0)), uD_t


class _Batch:
    """Concatenated device arrays of a batched solve."""

    def __init__(self, device):
        self.device = device
        self.inst = []
        self.s_e = []
        self.K0 = []
        self.dinv = []
        self.b = []
        self.node_off = self.elem_off = self.vec_off = 0
        self.ndofs = []

    def add(self, op, b_cols):
        nx, ny, nz = op.nx, op.ny, op.nz
        nrhs = len(b_cols)
        self.inst.append(H8Instance(nx, ny, nz, nrhs, self.node_off, self.elem_off, self.vec_off))
        self.s_e.append(op.s_e)
        self.K0.append(op.K0buf)
        self.dinv.append(op.dinv)
        self.b.extend(b_cols)
        self.ndofs.append((op.n_dof, nrhs))
        self.node_off += op.n_nodes
        self.elem_off += op.m
        self.vec_off += op.n_dof * nrhs

    def signature(self):
        """The batch composition: sorted (nx, ny, nz, nrhs) counts."""
        import collections
        c = collections.Counter((i.nx, i.ny, i.nz, i.nrhs) for i in self.inst)
        return ";".join("%dx%dx%dr%d*%d" % (k + (v,)) for k, v in sorted(c.items()))

    def solve(self, tol, maxiter, form=None):
        """form: _lib.BATCHED_* or None (the batched autotuner's choice)."""
        if form is None:
            form = choose_form(self)
        return self._solve(tol, maxiter, int(form))

    def _solve(self, tol, maxiter, form, check_rc=True):
        import torch
        n_inst = len(self.inst)
        arr = (H8Instance * n_inst)(*self.inst)
        s_e = torch.cat(self.s_e).contiguous()
        K0 = torch.cat(self.K0).contiguous()
        dinv = torch.cat(self.dinv).contiguous()
        b = torch.cat(self.b).contiguous()
        x = torch.empty_like(b)
        total_vec = b.numel()
        nbytes = _lib.lib().h8_batched_workspace_bytes(arr, n_inst, total_vec)
        if nbytes == 0:
            check(_lib.H8_ERR_ARG, "h8_batched_workspace_bytes")
        ws = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        n_sys = sum(n for _, n in self.ndofs)
        status = (C.c_int32 * n_sys)()
        iters = (C.c_int64 * n_sys)()
        rel = (C.c_double * n_sys)()
        if maxiter is None:
            maxiter = 10 * max(nd for nd, _ in self.ndofs)
        with torch.cuda.device(self.device):
            rc = _lib.lib().h8_pcg_batched_form_f64(arr, n_inst, self.node_off, self.elem_off, total_vec, ptr(s_e),
                                                    ptr(K0), 1, ptr(dinv), ptr(b), ptr(x), float(tol), int(maxiter),
                                                    ptr(ws), ws.numel(), status, iters, rel, int(form),
                                                    _lib.stream_handle(self.device))
        stats = {"status": np.array(status[:]), "iterations": np.array(iters[:]), "rel_residual": np.array(rel[:]),
                 "elapsed_ms": _lib.lib().h8_batched_last_ms(),
                 "launches": int(_lib.lib().h8_batched_last_launches()), "rc": rc, "form": FORM_NAMES[form]}
        if check_rc:
            check(rc)
        out, o = [], 0
        for nd, nrhs in self.ndofs:
            out.append(x[o:o + nd * nrhs].reshape(nrhs, nd).T)
            o += nd * nrhs
        return out, stats


FORM_NAMES = {_lib.BATCHED_LAUNCHED: "launched", _lib.BATCHED_PERSISTENT_8X16: "persistent-8x16",
              _lib.BATCHED_PERSISTENT_16X16: "persistent-16x16"}
# candidates in order of preference on near-ties: the one persistent launch
# first, then the launched form, then the narrow-column persistent form
CANDIDATES = (_lib.BATCHED_PERSISTENT_16X16, _lib.BATCHED_LAUNCHED, _lib.BATCHED_PERSISTENT_8X16)
TUNE_ITERS = 40
TIE = 0.02


def choose_form(batch, cache=True):
    """The batched autotuner: the fastest form for this batch's composition on
    this device, timed once (TUNE_ITERS iterations of every system at tol 0,
    the solve's own device time) and then served from the cache; the first of
    CANDIDATES within TIE of the fastest wins."""
    from . import autotune
    key = "%s|batched|%s" % (autotune._device_key(), batch.signature())
    path = autotune.cache_path()
    data = autotune._load(path) if cache else {}
    if key in data:
        return int(data[key][0])
    timed = []
    for form in CANDIDATES:
        batch._solve(0.0, TUNE_ITERS, form, check_rc=False)          # warm-up (graphs, code objects)
        _, st = batch._solve(0.0, TUNE_ITERS, form, check_rc=False)
        timed.append((form, st["elapsed_ms"]))
    fastest = min(t for _, t in timed)
    form, ms = next((f, t) for f, t in timed if t <= fastest * (1.0 + TIE))
    if cache:
        data = autotune._load(path)
        data[key] = [form, ms, {FORM_NAMES[f]: t for f, t in timed}]
        autotune._store(path, data)
    return form


def solve_static_batched(fesystem, F_batch, tol=1e-8, maxiter=None, return_stats=False, form=None):
    """U_batch (n_dof, n_batch) for the columns of F_batch (fe/batched_solve.py:105-164).
    form: _lib.BATCHED_* or None (the batched autotuner's choice)."""
    F_batch = np.asarray(F_batch, dtype=float)
    if F_batch.ndim != 2 or F_batch.shape[0] != fesystem.n_dof:
        raise ValueError(f"F_batch must have shape ({fesystem.n_dof}, n_batch), got {F_batch.shape}")
    fesystem._require_stiffness("solve_static_batched")
    op = fesystem._op
    op.set_free_mask(fesystem._free_mask())
    cols, uD = _dirichlet_columns(fesystem, F_batch)
    batch = _Batch(op.device)
    batch.add(op, cols)
    (X,), stats = batch.solve(tol, maxiter, form)
    if uD is not None:
        X = X + uD[:, None]
    U = X.cpu().numpy()
    return (U, stats) if return_stats else U


def solve_many(systems, tol=1e-8, maxiter=None, form=None):
    """systems: iterable of (FESystem, F_batch (n_dof_i, k_i)).  Returns
    ([U_i (n_dof_i, k_i) device tensors], stats).  Every FESystem must be
    assembled on the same device.  form: _lib.BATCHED_* or None (the batched
    autotuner's choice)."""
    import torch
    batch = None
    uDs = []
    flag = None
    for fs, F_batch in systems:
        fs._require_stiffness("solve_many")
        op = fs._op
        if flag is None:
            flag = torch.zeros(1, dtype=torch.int32, device=op.device)
        op.set_free_mask(fs._free_mask(), zero_diag_flag=flag)
        F_batch = np.asarray(F_batch, dtype=float)
        if F_batch.ndim == 1:
            F_batch = F_batch[:, None]
        cols, uD = _dirichlet_columns(fs, F_batch)
        if batch is None:
            batch = _Batch(op.device)
        elif op.device != batch.device:
            raise ValueError("solve_many: all systems must live on one GPU")
        batch.add(op, cols)
        uDs.append(uD)
    if batch is None:
        raise ValueError("solve_many: no systems given")
    if int(flag.item()):
        raise ValueError("jacobi_preconditioner: A has a zero diagonal entry.")
    X, stats = batch.solve(tol, maxiter, form)
    X = [x if uD is None else x + uD[:, None] for x, uD in zip(X, uDs)]
    return X, stats
