"""FEField: a full-length displacement vector with DOF names (the subset of
fe/fields.py:36-66 the static path returns)."""
from __future__ import annotations

import numpy as np


class FEField(np.ndarray):
    def __new__(cls, data, n_nodes=None, dofs_per_node=3, dof_names=("ux", "uy", "uz"), mesh=None,
                label=None):
        obj = np.asarray(data).view(cls)
        obj._n_nodes = n_nodes
        obj._npn = dofs_per_node
        obj._dof_names = tuple(dof_names) if dof_names is not None else None
        obj._mesh = mesh
        obj.label = label
        return obj

    def __array_finalize__(self, obj):
        if obj is None:
            return
        for name in ("_n_nodes", "_npn", "_dof_names", "_mesh", "label"):
            setattr(self, name, getattr(obj, name, None))

    def __array_wrap__(self, out_arr, context=None, return_scalar=False):
        if getattr(out_arr, "ndim", 1) == 0:
            return out_arr[()]
        return np.ndarray.__array_wrap__(self, out_arr, context)

    @property
    def nodal(self):
        """(n_nodes, dofs_per_node) view."""
        return np.asarray(self).reshape(-1, self._npn)

    def component(self, name, nodes=None):
        names = self._dof_names or ()
        idx = names.index(name) if isinstance(name, str) else int(name)
        vals = self.nodal[:, idx]
        if nodes is None:
            return vals
        if isinstance(nodes, str):
            nodes = self._mesh.node_set(nodes)
        return vals[np.asarray(nodes, dtype=int)]

    def magnitude(self):
        return np.linalg.norm(self.nodal, axis=1)
