"""ReducedSystem on the HIP backend: the constrained linear system over the kept
DOFs and the maps to and from the full DOF vector (fe/linear_system.py:26-140).

The reference returns the reduced matrix itself (sparse T^T K_ff T); this
backend is matrix-free, so ``K`` is a scipy LinearOperator whose product runs on
the GPU: scatter the kept entries into a full-length vector, expand to the slave
DOFs, apply the grid operator K, contract onto the masters, gather the kept
entries (one host-device round trip per product).  ``F``, ``free``, ``fixed``,
``u_fixed``, ``n_dof``, ``restrict``, ``expand``, ``recover`` and ``solve`` keep
the reference's meaning.
"""
from __future__ import annotations

import numpy as np


class ReducedSystem:
    def __init__(self, system, op, cm, w, b):
        import torch
        from scipy.sparse.linalg import LinearOperator
        self._system, self._op, self._cm = system, op, cm
        self.free = np.flatnonzero(cm.keep_mask).astype(np.int64)
        self.fixed = system.fixed_dofs_array
        self.u_fixed = system._fixed_dof_values_array(self.fixed)
        self.n_dof = system.n_dof
        self.slave = cm.slave
        self._w = w
        self._keep_t = torch.as_tensor(self.free, device=op.device)
        self.F = b[self._keep_t].cpu().numpy()
        n = len(self.free)
        self.K = LinearOperator((n, n), matvec=self._matvec, dtype=np.float64)

    @property
    def n_free(self):
        return len(self.free)

    def _matvec(self, x):
        import torch
        op, cm = self._op, self._cm
        v = torch.zeros(self.n_dof, dtype=torch.float64, device=op.device)
        v[self._keep_t] = torch.as_tensor(np.asarray(x, dtype=float).reshape(-1), device=op.device)
        cm.expand(v)
        y = cm.contract(op.matvec(v))
        return y[self._keep_t].cpu().numpy()

    def restrict(self, v):
        """Full vector -> kept-DOF part."""
        return np.asarray(v)[self.free]

    def reduce_matrix(self, A):
        raise NotImplementedError("reduce_matrix: backend='hip' holds no assembled matrices")

    def expand(self, u_free, include_prescribed=True):
        """Kept-DOF vector -> full-length array: the slave DOFs from the
        constraints, the prescribed values at fixed DOFs (include_prescribed)."""
        u_free = np.asarray(u_free, dtype=float)
        if u_free.shape != (self.n_free,):
            raise ValueError(f"expected {self.n_free} free-DOF rows, got {u_free.shape[0]}")
        full = np.zeros(self.n_dof)
        full[self.free] = u_free
        cm = self._cm
        if len(cm.slave):
            s_ptr, s_master, s_coef = (t.cpu().numpy() for t in cm._t[1:4])
            owner = np.repeat(np.arange(len(cm.slave)), np.diff(s_ptr))   # slave of each (slave, master) entry
            dep = np.bincount(owner, weights=s_coef * full[s_master], minlength=len(cm.slave))
            full[cm.slave] = dep + (cm.g if include_prescribed else 0.0)
        if include_prescribed:
            full[self.fixed] = self.u_fixed
        return full

    def recover(self, u_free, label="displacement", include_prescribed=True):
        return self._system.field(self.expand(u_free, include_prescribed), label=label)

    def solve(self, label="displacement", tol=1e-8, max_iter=None):
        """Jacobi-PCG of the eliminated system on the GPU, recovered to the full U."""
        u = self._system._solve_constrained(tol, max_iter, self._system.F)
        return self._system.field(u.cpu().numpy(), label=label)
