"""Linear multi-point constraints on the HIP backend (fe/solver.py:1341-1479,
fe/boundary.py:98-166).

``reduce_constraints`` eliminates the constraints sum_j c_j u_j = value into
slave DOFs u_s = g_s - sum_m mat[s][m] u_m over kept (master) DOFs with the
reference's rule: Gauss-Jordan elimination, pivoting on the largest remaining
|coefficient| (first in row-major order over the sorted DOFs), prescribed DOFs
moved to the right-hand side, leftover rows checked for consistency.  Rows that
share no DOF never interact in that elimination, so each connected group of
rows is reduced on its own -- the same pivots and values as eliminating them
all at once (the tolerances are scaled by the group's largest entry instead
of the whole set's), at the cost of the largest group (a face tie is
thousands of independent pairs).

``ConstraintMap`` holds the reduced map on the device in the layout of
``h8_cmap`` (include/h8pcg.h): T_sm = -mat[s][m], by slave and by master.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, ptr


def _reduce_group(rows, rhs, tol):
    """Eliminate one connected group of constraint rows ({dof: coefficient},
    right-hand side): Gauss-Jordan on the augmented table with, at every step,
    the entry of largest magnitude among the rows and columns not yet used as
    the pivot -- the elimination order of fe/boundary.py:98-166, so the same
    DOFs become slaves.  Returns (slaves, {slave: {master: coefficient}},
    {slave: value})."""
    dofs = sorted({dof for row in rows for dof in row})
    if not dofs:
        for v in rhs:
            if abs(v) > tol * max(1.0, abs(v)):
                raise ValueError("inconsistent constraint: every DOF it involves is prescribed and the "
                                 "prescribed values do not satisfy it.")
        return [], {}, {}
    col_of = {dof: j for j, dof in enumerate(dofs)}
    ncol = len(dofs)
    table = np.zeros((len(rows), ncol + 1))
    for i, row in enumerate(rows):
        for dof, coef in row.items():
            table[i, col_of[dof]] = coef
        table[i, ncol] = rhs[i]
    magnitude = max(1.0, np.abs(table).max())
    pivot_row, pivot_col = [], []
    for _ in range(min(len(rows), ncol)):
        cand = np.abs(table[:, :ncol])
        cand[pivot_row, :] = 0.0
        cand[:, pivot_col] = 0.0
        i, j = np.unravel_index(np.argmax(cand), cand.shape)
        if cand[i, j] <= tol * magnitude:
            break
        table[i] /= table[i, j]
        others = np.flatnonzero((np.arange(len(rows)) != i) & (table[:, j] != 0.0))
        for r in others:
            table[r] -= table[r, j] * table[i]
        pivot_row.append(i)
        pivot_col.append(j)
    leftover = [r for r in range(len(rows)) if r not in pivot_row]
    if any(abs(table[r, ncol]) > 1e-8 * magnitude for r in leftover):
        raise ValueError("inconsistent linear constraints: the equations contradict one another.")
    masters = [j for j in range(ncol) if j not in pivot_col]
    slaves, coeffs, values = [], {}, {}
    for i, j in zip(pivot_row, pivot_col):
        s = dofs[j]
        slaves.append(s)
        coeffs[s] = {dofs[k]: table[i, k] for k in masters if abs(table[i, k]) > 1e-14}
        values[s] = float(table[i, ncol])
    return slaves, coeffs, values


def reduce_constraints(rows, values, fixed_values, tol=1e-10):
    """(slave DOFs, {slave: {master: coefficient}}, {slave: g}) with
    u_slave = g_slave - sum coefficient * u_master (fe/boundary.py:98-166)."""
    clean, rhs = [], []
    for r, v in zip(rows, values):
        row, v = {}, float(v)
        for d, c in r.items():
            if d in fixed_values:
                v -= c * fixed_values[d]
            else:
                row[d] = row.get(d, 0.0) + c
        clean.append(row)
        rhs.append(v)
    # connected groups of rows (rows sharing a DOF), by union-find over DOFs
    parent = {}

    def find(x):
        while parent.setdefault(x, x) != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for i, r in enumerate(clean):
        parent.setdefault(("row", i), ("row", i))
        for d in r:
            a, b = find(("row", i)), find(("dof", d))
            if a != b:
                parent[b] = a
    groups = {}
    for i in range(len(clean)):
        groups.setdefault(find(("row", i)), []).append(i)
    slave, mat, g = [], {}, {}
    for idx in groups.values():
        s_, m_, g_ = _reduce_group([clean[i] for i in idx], [rhs[i] for i in idx], tol)
        slave.extend(s_)
        mat.update(m_)
        g.update(g_)
    return np.array(slave, dtype=np.int64), mat, g


class ConstraintMap:
    """The reduced map on the device (h8_cmap): T_sm = -mat[s][m] by slave and
    by master, the slave offsets g, and the ctypes struct pointing at them."""

    def __init__(self, slave, mat, g, device):
        import torch
        self.slave = np.asarray(slave, dtype=np.int64)
        self.g = np.array([g[int(s)] for s in self.slave], dtype=float)
        s_ptr, s_master, s_coef = [0], [], []
        by_master = {}
        for s in self.slave:
            for m, c in sorted(mat[int(s)].items()):
                s_master.append(int(m))
                s_coef.append(-float(c))
                by_master.setdefault(int(m), []).append((int(s), -float(c)))
            s_ptr.append(len(s_master))
        self.master = np.array(sorted(by_master), dtype=np.int64)
        m_ptr, m_slave, m_coef = [0], [], []
        for m in self.master:
            for s, c in by_master[int(m)]:
                m_slave.append(s)
                m_coef.append(c)
            m_ptr.append(len(m_slave))
        if len(s_master) >= 2 ** 31 or len(m_slave) >= 2 ** 31:
            raise ValueError("constraint map too large")

        def dev(a, dt):
            return torch.as_tensor(np.asarray(a, dtype=dt), device=device).contiguous()

        self._t = [dev(self.slave, np.int64), dev(s_ptr, np.int32), dev(s_master, np.int64), dev(s_coef, float),
                   dev(self.master, np.int64), dev(m_ptr, np.int32), dev(m_slave, np.int64), dev(m_coef, float)]
        t = self._t
        self.c = _lib.H8CMap(len(self.slave), ptr(t[0]), ptr(t[1]), ptr(t[2]), ptr(t[3]), len(self.master),
                             ptr(t[4]), ptr(t[5]), ptr(t[6]), ptr(t[7]))

    def expand(self, v):
        """v_s = sum_m T_sm v_m at every slave (in place, device tensor)."""
        check(_lib.lib().h8_constraint_expand_f64(C.byref(self.c), ptr(v), _lib.stream_handle(v.device)),
              "h8_constraint_expand_f64")
        return v

    def contract(self, v):
        """v_m += sum_s T_sm v_s at every master (in place, device tensor)."""
        check(_lib.lib().h8_constraint_contract_f64(C.byref(self.c), ptr(v), _lib.stream_handle(v.device)),
              "h8_constraint_contract_f64")
        return v
