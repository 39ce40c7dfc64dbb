"""FESystem on the HIP backend: the drop-in for fea_engine's linear-elastic H8
assemble/solve path (fe/solver.py FESystem).

Same constructor, same ``assemble_stiffness`` / ``fix_dofs`` / ``add_nodal_force``
/ ``add_consistent_facet_load`` / ``free_dofs`` / ``solve_static`` methods,
argument meanings and exception types as the reference (fe/solver.py:58-146,
447-518, 569-697, 1109-1270, 1399-1509, 1726-2045).  What changes is where
the work happens: assembly stores one 24x24 element matrix K0 (computed on the
GPU) and a per-element scale s_e instead of a global K, and solve_static runs
the matrix-free Jacobi-PCG in libh8pcg.so.  ``backend="hip"`` is the only
backend of this package.
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _lib
from .coefficients import Coefficient, factor_scales, has_coefficient
from .elements import Hex8Solid3D
from .facet_loads import face_load_shares_batched
from .fields import FEField
from .mesh import structured_grid


def _integer_ids(values, what, who):
    """A flat int array from ints, integral floats, sets or arrays; ValueError otherwise."""
    if isinstance(values, (set, frozenset)):
        values = sorted(values)
    arr = np.asarray(values).ravel()
    if arr.size == 0 or arr.dtype.kind in "iu":
        return arr.astype(np.int64)
    if arr.dtype.kind == "f" and np.array_equal(arr, np.floor(arr)):
        return arr.astype(np.int64)
    raise ValueError(f"{who}: {what} have to be whole numbers (got dtype {arr.dtype}).")


def _node_selection(mesh, spec, who):
    """Node ids of `spec`: ids, or one or more node-set names of `mesh`."""
    names = [spec] if isinstance(spec, str) else None
    if names is None and isinstance(spec, (list, tuple)) and spec and all(isinstance(v, str) for v in spec):
        names = list(spec)
    if names is not None:
        sets = getattr(mesh, "node_sets", {})
        absent = [nm for nm in names if nm not in sets]
        if absent:
            raise ValueError(f"{who}: unknown node set {absent} (the mesh defines {sorted(sets)}); create it "
                             f"first with mesh.add_node_set or mesh.select_nodes(..., name=...).")
        spec = np.unique(np.concatenate([np.asarray(sets[nm], dtype=np.int64) for nm in names]))
    ids = _integer_ids(spec, "node ids", who)
    if ids.size == 0:
        raise ValueError(f"{who}: no nodes selected -- the id list is empty, which is what a coordinate "
                         f"selector returns when it matched no node (check its value and tol).")
    n = len(mesh.nodes)
    outside = np.unique(ids[(ids < 0) | (ids >= n)])
    if outside.size:
        raise ValueError(f"{who}: node ids {outside[:5].tolist()} are outside the mesh, whose ids run "
                         f"0..{n - 1}.")
    return ids


def _component_selection(spec, npn, by_name, who):
    """Local component indices of `spec`: an index, a name, or a collection of
    either; each must lie in 0..npn-1."""
    if isinstance(spec, (set, frozenset)):
        items = sorted(spec, key=str)
    elif isinstance(spec, str) or np.ndim(spec) == 0:
        items = [spec]
    else:
        items = list(spec)
    comps = _integer_ids([by_name(v) if isinstance(v, str) else v for v in items], "component indices", who)
    if comps.size == 0:
        raise ValueError(f"{who}: no DOF indices given.")
    wrong = sorted(set(comps[(comps < 0) | (comps >= npn)].tolist()))
    if wrong:
        raise ValueError(f"{who}: component index {wrong} does not exist -- Hex8Solid3D has {npn} DOF(s) per "
                         f"node, numbered 0..{npn - 1}.")
    return comps


class FESystem:
    def __init__(self, mesh, elem_formulation, thickness=1.0, sparse=False, backend="hip", device="cuda"):
        if backend != "hip":
            raise ValueError(
                f"FESystem: unknown backend={backend!r} -- this package provides backend='hip' "
                f"(matrix-free Jacobi-PCG on an AMD GPU); the reference's 'scipy', 'torch' and "
                f"'auto' backends live in fea_engine itself.")
        from .hex20 import Hex20Solid3D
        if not isinstance(elem_formulation, (Hex8Solid3D, Hex20Solid3D)):
            raise NotImplementedError(
                f"FESystem(backend='hip') supports Hex8Solid3D, Hex8SolidBbar and Hex20Solid3D, got "
                f"{type(elem_formulation).__name__}")
        self._hex20 = isinstance(elem_formulation, Hex20Solid3D)
        _lib.lib()   # fail fast when the device library is missing (as backend='torch' does for torch)
        self.backend = backend
        self.device = device
        self.mesh = mesh
        self.elem = elem_formulation
        self.thickness = thickness
        self.sparse = sparse
        self.npn = 3
        self.n_dof = 3 * len(mesh.nodes)
        self.F = np.zeros(self.n_dof)
        self.fixed_dofs = set()
        self.fixed_dof_values = {}
        self._assembled = set()
        self._warned_reassembly = set()
        self._op = None
        self._D_ref = None
        self._K_cache = None
        self._structure = None
        self._constraints = ()
        self.last_solve_stats = None

    # ---- argument resolution (the contract of fe/solver.py:447-496) -------------------------
    # Node arguments: integer ids, or the name(s) of node sets registered on the
    # mesh.  Component arguments: local indices or names ("uy").  Both raise
    # ValueError; the message fragments the reference's tests match on are kept.
    @property
    def dof_names(self):
        return self.elem.local_dof_names()

    def dof_index(self, name):
        return self._dof_index_by_name(name, "dof_index")

    def _dof_index_by_name(self, name, who):
        try:
            return self.elem.dof_index(name)
        except ValueError as err:
            raise ValueError(f"{who}: {err}") from None

    def _check_nodes(self, node_ids, who):
        return _node_selection(self.mesh, node_ids, who)

    def _check_local_dofs(self, dof_indices, who):
        return _component_selection(dof_indices, self.npn, lambda nm: self._dof_index_by_name(nm, who), who)

    # ---- boundary conditions and loads (fe/solver.py:1109-1270) ----------------------------
    def fix_dofs(self, node_ids, dof_indices, value=0.0):
        nodes = self._check_nodes(node_ids, "fix_dofs")
        comps = self._check_local_dofs(dof_indices, "fix_dofs")
        gdofs = (self.npn * nodes[:, None] + comps[None, :]).ravel().tolist()
        self.fixed_dofs.update(gdofs)
        self.fixed_dof_values.update(dict.fromkeys(gdofs, float(value)))

    def add_nodal_force(self, node_ids, dof_index, total_force):
        nodes = self._check_nodes(node_ids, "add_nodal_force")
        dof = int(self._check_local_dofs(dof_index, "add_nodal_force")[0])
        share = total_force / len(nodes)
        np.add.at(self.F, self.npn * nodes + dof, share)

    def add_consistent_facet_load(self, formulation, facets, dof_index, traction, quad_order=2, thickness=1.0):
        facets = [tuple(f) for f in facets]
        self._check_nodes([n for f in facets for n in f], "add_consistent_facet_load")
        dof_index = int(self._check_local_dofs(dof_index, "add_consistent_facet_load")[0])
        if not facets:
            return
        if self._hex20:
            raise NotImplementedError("add_consistent_facet_load with Hex20Solid3D (8-node faces) is not "
                                      "available on backend='hip'; apply nodal forces (add_nodal_force).")
        if any(len(f) != 4 for f in facets):
            raise ValueError("add_consistent_facet_load: Hex8 faces have 4 nodes")
        ids = np.asarray(facets, dtype=int)
        shares = face_load_shares_batched(np.asarray(self.mesh.nodes)[ids], traction * thickness,
                                          quad_order=quad_order)
        np.add.at(self.F, self.npn * ids.ravel() + dof_index, shares.ravel())

    # ---- DOF partition (fe/solver.py:1399-1401,1481-1494) ----------------------------------
    @property
    def free_dofs(self):
        mask = np.ones(self.n_dof, dtype=bool)
        if self.fixed_dofs:
            mask[np.fromiter(self.fixed_dofs, dtype=np.int64)] = False
        return np.nonzero(mask)[0]

    @property
    def fixed_dofs_array(self):
        return np.array(sorted(self.fixed_dofs), dtype=int)

    def _fixed_dof_values_array(self, fixed):
        return np.array([self.fixed_dof_values.get(int(d), 0.0) for d in fixed])

    def _free_mask(self):
        mask = np.ones(self.n_dof, dtype=np.uint8)
        if self.fixed_dofs:
            mask[np.fromiter(self.fixed_dofs, dtype=np.int64)] = 0
        return mask

    # ---- assembly (fe/solver.py:498-518,569-697) -------------------------------------------
    def _begin_assembly(self, accumulate, who):
        # a repeated assemble_stiffness replaces K (accumulate=True adds to it);
        # the first replacement warns once
        if accumulate or "K" not in self._assembled or "K" in self._warned_reassembly:
            return
        self._warned_reassembly.add("K")
        warnings.warn(f"{who}: K was already assembled; this call replaces it (pass accumulate=True to add "
                      f"the new contribution instead).", UserWarning, stacklevel=3)

    def assemble_stiffness(self, D, method="full", vectorized=False, accumulate=False, **kwargs):
        """Stores K0 (24x24, device) and s_e (per element, device).

        D: a 6x6 matrix, or coefficients.by_element([...6x6...]) for per-element
        material (the matrices must be scalar multiples of one another).
        kwargs: thickness= (scales every element), gauss_order= (2 only),
        element_scale= (an (m,) array multiplying each element, e.g. a SIMP scale)."""
        known = ("full", "reduced", "hourglass_stabilized")
        if method not in known:
            raise ValueError(f"assemble_stiffness: unknown method {method!r}; the integration methods are "
                             f"{', '.join(map(repr, known))}.")
        if vectorized and method != "full":
            raise ValueError(f"assemble_stiffness: vectorized=True only supports method='full', and "
                             f"method={method!r} was asked for.")
        if method != "full":
            raise NotImplementedError(f"assemble_stiffness(method={method!r}) is not available on backend='hip'")
        kw = dict(kwargs)
        thickness = float(kw.pop("thickness", 1.0))
        gauss_order = kw.pop("gauss_order", None)
        elem_scale = kw.pop("element_scale", None)
        kw.pop("chunk_size", None)
        kw.pop("backend", None)
        kw.pop("device", None)
        if kw:
            raise TypeError(f"assemble_stiffness: unsupported keyword(s) {sorted(kw)}")
        if self._hex20:
            if gauss_order not in (None, 3):
                raise NotImplementedError("backend='hip' integrates Hex20 with the 3x3x3 rule only")
        elif gauss_order not in (None, 2):
            raise NotImplementedError("backend='hip' integrates Hex8 with the 2x2x2 rule only")
        self._begin_assembly(accumulate, "assemble_stiffness")
        m = len(self.mesh.elements)
        if has_coefficient(D):
            D_ref, s = factor_scales(D)
            if len(s) != m:
                raise ValueError(f"Coefficient.by_element: {len(s)} values but the mesh has {m} elements.")
        else:
            D_ref = np.asarray(D, dtype=float)
            if D_ref.shape != (6, 6):
                raise ValueError(f"assemble_stiffness: D must be 6x6 for Hex8Solid3D, got {D_ref.shape}")
            s = np.ones(m)
        s = s * thickness
        if elem_scale is not None:
            es = np.asarray(elem_scale, dtype=float).reshape(-1)
            if es.shape != (m,):
                raise ValueError(f"element_scale must have one entry per element ({m})")
            s = s * es
        if self._hex20:
            from .hex20 import structured_hex20, Hex20Operator
            st = self._structure if self._structure is not None else structured_hex20(self.mesh)
            if st is None:
                raise NotImplementedError(
                    "FESystem(backend='hip') with Hex20Solid3D needs a structured box mesh "
                    "(box_mesh_hex20 connectivity with congruent bricks)")
            operator_cls = Hex20Operator
        else:
            st = self._structure if self._structure is not None else structured_grid(self.mesh)
            if st is None:
                raise NotImplementedError(
                    "FESystem(backend='hip'): the matrix-free operator needs a structured box mesh "
                    "(box_mesh connectivity with congruent bricks)")
            from .operator import GridOperator as operator_cls
        self._structure = st
        nx, ny, nz, coords0 = st
        if accumulate and self._op is not None:
            ratio = _proportional(self._D_ref, D_ref)
            s = self._op.s_e.cpu().numpy() + ratio * s
            self._op.set_scale(s)
        else:
            ke0 = self.elem.stiffness_batched(coords0[None], D_ref, device=self.device)[0]
            self._op = operator_cls(nx, ny, nz, ke0, s, device=self.device)
            self._D_ref = D_ref
        self._assembled.add("K")
        self._K_cache = None
        self._stiffness_args = (D, dict(kwargs))

    def _require_stiffness(self, who):
        if self._op is not None:
            return
        raise RuntimeError(f"{who}: nothing is assembled yet (K is all zeros); run assemble_stiffness(...) "
                           f"first.")

    @property
    def operator(self):
        self._require_stiffness("operator")
        return self._op

    @property
    def K(self):
        """The assembled global stiffness as scipy CSR, built on demand on the GPU
        (COO values s_e*K0 summed by torch's coalesce) and copied to the host."""
        self._require_stiffness("K")
        if self._K_cache is None and self._hex20:
            self._K_cache = self._op.csr(np.asarray(self.mesh.elements))
        if self._K_cache is None:
            self._K_cache = _materialize_csr(self._op, np.asarray(self.mesh.elements))
        return self._K_cache

    # ---- solve (fe/solver.py:1726-2045) ----------------------------------------------------
    def _dirichlet_rhs(self, op, F_host):
        """b = mask(F - K u_D) on the device and the full u_D vector (None if all zero).
        F_host: a host array, or a float64 tensor already on the operator's device
        (used in place: no host round trip)."""
        import torch
        fixed = self.fixed_dofs_array
        u_fixed = self._fixed_dof_values_array(fixed)
        if isinstance(F_host, torch.Tensor) and F_host.device == torch.device(op.device):
            if F_host.dtype != torch.float64 or F_host.numel() != self.n_dof:
                raise ValueError(f"F must be a float64 vector of {self.n_dof} entries")
            F_t = F_host.reshape(-1)
        else:
            F_t = torch.as_tensor(np.asarray(F_host, dtype=float), device=op.device)
        if len(fixed) and np.any(u_fixed):
            uD = np.zeros(self.n_dof)
            uD[fixed] = u_fixed
            uD_t = torch.as_tensor(uD, device=op.device)
            return op.masked(F_t, op.matvec(uD_t)), uD_t
        return op.masked(F_t), None

    def _rigid_null_basis(self):
        """Orthonormal basis (n_dof, r) of the rigid-body motions that vanish at every
        fixed DOF -- the null space of Kff for a connected mesh with positive element
        scales -- or None when the supports remove all six.  fe/solver.py:1583-1678
        finds the same directions as near-zero eigenvalues of a dense Kff; here they
        follow from the node coordinates: translations and rotations about the
        centroid, combined to vanish on the fixed set."""
        fixed = self.fixed_dofs_array
        key = (len(fixed), hash(fixed.tobytes()))
        cache = getattr(self, "_rigid_cache", None)
        if cache is not None and cache[0] == key:
            return cache[1]
        X = np.asarray(self.mesh.nodes, dtype=float)
        Y = X - X.mean(axis=0)
        Y /= max(float(np.ptp(X, axis=0).max()), 1e-300)

        def rig(dofs):
            a, c = dofs // 3, dofs % 3
            y = Y[a]
            R = np.zeros((len(dofs), 6))
            R[np.arange(len(dofs)), c] = 1.0
            R[:, 3] = np.where(c == 1, -y[:, 2], np.where(c == 2, y[:, 1], 0.0))
            R[:, 4] = np.where(c == 0, y[:, 2], np.where(c == 2, -y[:, 0], 0.0))
            R[:, 5] = np.where(c == 0, -y[:, 1], np.where(c == 1, y[:, 0], 0.0))
            return R

        if len(fixed) == 0:
            N = np.eye(6)
        else:
            _, s, Vt = np.linalg.svd(rig(fixed), full_matrices=True)
            rank = int(np.sum(s > 1e-10 * max(float(s[0]), 1.0)))
            N = Vt[rank:].T
        Q = None
        if N.shape[1] > 0:
            M = rig(np.arange(self.n_dof)) @ N
            M[fixed] = 0.0
            Q, _ = np.linalg.qr(M)
            Q[fixed] = 0.0      # zero rows stay zero (Householder leaves roundoff there)
        self._rigid_cache = (key, Q)
        return Q

    # ---- linear constraints (fe/solver.py:1341-1397) -----------------------------------------
    def add_constraint(self, terms, value=0.0):
        """General linear constraint sum_i c_i u_i = value; terms: (node, dof, coefficient).
        Eliminated exactly in the solve (fe/solver.py:1341-1372)."""
        if not terms:
            raise ValueError("add_constraint: no terms given.")
        dofs, coefs = [], []
        for node, dof, c in terms:
            n = self._check_nodes([node] if not isinstance(node, str) else node, "add_constraint")
            if len(n) != 1:
                raise ValueError("add_constraint: each term needs exactly one node.")
            d = self._check_local_dofs(dof, "add_constraint")
            if len(d) != 1:
                raise ValueError("add_constraint: each term needs exactly one DOF.")
            dofs.append(3 * int(n[0]) + int(d[0]))
            coefs.append(float(c))
        if len(set(dofs)) != len(dofs):
            raise ValueError("add_constraint: a DOF appears twice in one constraint; merge its terms.")
        if not np.any(np.asarray(coefs)):
            raise ValueError("add_constraint: all coefficients are zero.")
        self._constraints = tuple(self._constraints) + ((np.array(dofs), np.array(coefs), float(value)),)

    def tie(self, slave_nodes, master_nodes, dof_indices, offset=None, tol=1e-8):
        """u_slave = u_master for each slave node and the master node at x_slave +
        offset (default: coincident nodes), fe/solver.py:1374-1393."""
        from scipy.spatial import cKDTree
        s = self._check_nodes(slave_nodes, "tie")
        m = self._check_nodes(master_nodes, "tie")
        dofs = self._check_local_dofs(dof_indices, "tie")
        nodes = np.asarray(self.mesh.nodes, dtype=float)
        off = np.zeros(nodes.shape[1]) if offset is None else np.asarray(offset, dtype=float)
        dist, idx = cKDTree(nodes[m]).query(nodes[s] + off)
        if np.any(dist > tol):
            bad = s[np.argmax(dist)]
            raise ValueError(f"tie: no master node within {tol:g} of slave node {int(bad)} (offset {off}).")
        for sn, j in zip(s, idx):
            mn = int(m[j])
            if mn == int(sn):
                continue
            for d in dofs:
                self.add_constraint([(int(sn), int(d), 1.0), (mn, int(d), -1.0)])

    def clear_constraints(self):
        """Remove all linear constraints added with add_constraint / tie."""
        self._constraints = ()

    def _constrained_setup(self, F_host):
        """Device set-up of the eliminated system (fe/solver.py:1403-1479): the
        constraints reduced as fe/boundary.py:98-166 does (ConstraintMap, empty
        without constraints), dinv of T^T K T (zero at fixed and slave DOFs),
        w = prescribed values and slave offsets, and b = mask(T^T (F - K w)).
        Returns (op, cm, w_t, b)."""
        import ctypes as C
        import torch
        from .constraints import ConstraintMap, reduce_constraints
        op = self._op
        fixed = self.fixed_dofs_array
        u_fixed = self._fixed_dof_values_array(fixed)
        rows = [dict(zip(d.tolist(), c.tolist())) for d, c, _ in self._constraints]
        vals = [v for _, _, v in self._constraints]
        slave, mat, g = reduce_constraints(rows, vals, dict(zip(fixed.tolist(), u_fixed.tolist())))
        cm = ConstraintMap(slave, mat, g, op.device)
        mask = self._free_mask()
        mask[cm.slave] = 0
        op.set_free_mask(mask)
        with torch.cuda.device(op.device):
            _lib.check(_lib.lib().h8_constraint_diag_f64(C.byref(op.grid), _lib.ptr(op.s_e), _lib.ptr(op.K0buf),
                                                         C.byref(cm.c), _lib.ptr(op.dinv), op._stream()),
                       "h8_constraint_diag_f64")
            if len(cm.master) and float(op.dinv[torch.as_tensor(cm.master, device=op.device)].min()) <= 0.0:
                raise ValueError("jacobi_preconditioner: the constrained system has a zero diagonal entry.")
            w = np.zeros(self.n_dof)
            w[fixed] = u_fixed
            w[cm.slave] = cm.g
            w_t = torch.as_tensor(w, device=op.device)
            F_t = torch.as_tensor(np.asarray(F_host, dtype=float), device=op.device)
            b = (F_t - op.matvec(w_t)) if np.any(w) else F_t.clone()
            b = op.masked(cm.contract(b.contiguous()))
        cm.keep_mask = mask
        return op, cm, w_t, b

    def _solve_constrained(self, tol, maxiter, F_host):
        """Jacobi-PCG on T^T K T over the kept DOFs (h8_pcg_constrained_f64);
        u = T x + w."""
        import ctypes as C
        import torch
        op, cm, w_t, b = self._constrained_setup(F_host)
        if maxiter is None:
            maxiter = 10 * int(cm.keep_mask.sum())
        with torch.cuda.device(op.device):
            x = torch.zeros(self.n_dof, dtype=torch.float64, device=op.device)
            st = _lib.H8PcgStats()
            ws = op.workspace(64)
            rc = _lib.lib().h8_pcg_constrained_f64(C.byref(op.grid), _lib.ptr(op.s_e), _lib.ptr(op.K0buf),
                                                   _lib.ptr(op.dinv), C.byref(cm.c), _lib.ptr(b), _lib.ptr(x),
                                                   float(tol), int(maxiter), _lib.ptr(ws), ws.numel(),
                                                   C.byref(st), op._stream())
            stats = st.as_dict()
            stats["rc"] = rc
            _lib.check(rc)
            cm.expand(x)
        stats.update({"method": "pcg", "precision": "fp64", "constraints": len(self._constraints),
                      "slaves": int(len(cm.slave))})
        self.last_solve_stats = stats
        return x + w_t

    def form_linear_system(self, F=None, K=None):
        """The eliminated system K_r u_r = F_r over the kept DOFs as a
        ReducedSystem (fe/solver.py:1403-1479): K_r is a scipy LinearOperator
        whose products run on the GPU (T^T K T, matrix-free), F_r the reduced
        load; .recover(u_r) gives the full U.  "Bring your own solver":
        rs.recover(scipy.sparse.linalg.cg(rs.K, rs.F)[0]).  K= (another full
        matrix) is not available on the matrix-free backend."""
        from .linear_system import ReducedSystem
        if K is not None:
            raise NotImplementedError("form_linear_system(K=...): backend='hip' is matrix-free (K = sum_e s_e K0)")
        self._require_stiffness("form_linear_system")
        Fv = self.F if F is None else np.asarray(F, dtype=float)
        if Fv.shape != (self.n_dof,):
            raise ValueError(f"F must have shape ({self.n_dof},), got {Fv.shape}")
        op, cm, w_t, b = self._constrained_setup(Fv)
        return ReducedSystem(self, op, cm, w_t.cpu().numpy(), b)

    def solve_static_device(self, method=None, tol=1e-8, maxiter=None, x0=None, preconditioner=None,
                            precision="fp64", delta=None, F=None):
        """solve_static, returning the full displacement as a device tensor (no host copy).
        F: the load vector (default: the assembled one); a float64 tensor on the
        solve's device is used where it lies, without a host copy.
        preconditioner: None (the default: 'jacobi' in FP64, the reference's
        choice; 'mg' with precision='mixed'), 'jacobi', or 'mg' (geometric
        multigrid V-cycle, precision='mixed' only)."""
        self._require_stiffness("solve_static")
        if self._hex20 and (precision != "fp64" or preconditioner == "mg" or self._constraints):
            raise NotImplementedError("Hex20Solid3D on backend='hip' solves with the FP64 Jacobi-PCG "
                                      "(precision='fp64', no linear constraints).")
        if preconditioner is None:
            preconditioner = "mg" if precision == "mixed" and method in (None, "pcg") else "jacobi"
        if self._constraints:
            if preconditioner == "mg" and precision == "mixed":
                raise NotImplementedError("solve_static with linear constraints supports the FP64 Jacobi-PCG "
                                          "(precision='fp64') on backend='hip'.")
            if method not in (None, "pcg") or preconditioner != "jacobi" or precision != "fp64" or x0 is not None:
                raise NotImplementedError("solve_static with linear constraints supports the FP64 Jacobi-PCG "
                                          "(method None or 'pcg', no x0) on backend='hip'.")
            return self._solve_constrained(tol, maxiter, self.F if F is None else F)
        if method is None:
            method = "pcg"
        if method not in ("cg", "pcg"):
            raise ValueError(
                f"solve_static: unknown method={method!r} -- choose from 'cg', 'pcg' (the HIP backend "
                f"solves iteratively; the direct factorisations are CPU-only in the reference).")
        if method == "cg":
            preconditioner = None
        elif preconditioner in ("ssor", "ic0"):
            raise NotImplementedError(
                f"solve_static(method='pcg', preconditioner={preconditioner!r}) is not supported with "
                f"backend='hip'; use preconditioner='jacobi'.")
        elif preconditioner == "mg":
            if precision != "mixed":
                raise NotImplementedError("solve_static(preconditioner='mg') is the preconditioner of the FP32 "
                                          "iterations: use precision='mixed' (the FP64 solve is the reference's "
                                          "Jacobi-PCG).")
        elif preconditioner != "jacobi":
            raise ValueError(
                f"solve_static(method='pcg'): unknown preconditioner="
                f"{preconditioner!r} -- choose 'jacobi', 'ssor', 'ic0' or 'mg'.")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        op = self._op
        op.set_free_mask(self._free_mask(), preconditioner="jacobi" if preconditioner else None)
        mg = preconditioner == "mg"
        b, uD = self._dirichlet_rhs(op, self.F if F is None else F)
        # Supports that leave rigid-body motions free make Kff singular.  As in
        # fe/solver.py:1645-1678: a load with a component along them (beyond 1e-6 of
        # max|Ff|) is an under-constrained model; otherwise the result is the
        # minimum-norm solution -- that component is dropped from the load (so the
        # system is consistent) and from the PCG iterate.
        Q = self._rigid_null_basis()
        if Q is not None:
            import torch
            Q = torch.as_tensor(Q, device=b.device)
            proj = Q.T @ b
            bmax = float(b.abs().max())
            if bmax > 0 and float(proj.abs().max()) > 1e-6 * bmax:
                raise np.linalg.LinAlgError(
                    f"solve_static: the free-DOF stiffness matrix is singular ({Q.shape[1]} rigid-body "
                    f"mode(s) left free by the supports) and the load is not self-equilibrated against "
                    f"them -- an under-constrained model (likely a missing boundary condition). Add "
                    f"supports, or balance the load to zero net force and moment.")
            b = b - Q @ proj
            mg = False            # the coarse operators would be singular too
        if mg and not op.mg_available():
            mg = False            # the grid does not coarsen far enough
        n_free = self.n_dof - len(self.fixed_dofs)
        if maxiter is None:
            maxiter = 10 * n_free
        x0_t = None
        if x0 is not None:
            x0 = np.asarray(x0, dtype=float)
            full = np.zeros(self.n_dof)
            if x0.shape == (n_free,):
                full[self.free_dofs] = x0
            elif x0.shape == (self.n_dof,):
                full[self.free_dofs] = x0[self.free_dofs]
            else:
                raise ValueError(f"x0 must have {n_free} (free) or {self.n_dof} entries")
            x0_t = full
        if precision == "fp64":
            x, stats = op.pcg(b, tol=tol, maxiter=maxiter, x0=x0_t)
        else:
            x, stats = op.refined(b, tol=tol, maxiter=maxiter, delta=delta, x0=x0_t,
                                  preconditioner="mg" if mg else "jacobi")
        if Q is not None:
            x = x - Q @ (Q.T @ x)
        if uD is not None:
            x = x + uD
        stats["method"] = method
        stats["precision"] = precision
        self.last_solve_stats = stats
        return x

    def solve_static(self, verbose=False, method=None, **method_kwargs):
        """Jacobi-PCG on the GPU (method None, 'pcg' or 'cg'); returns the full-length
        displacement as an FEField on the host.  method_kwargs: tol (1e-8),
        maxiter (10 * free DOFs), x0, preconditioner ('jacobi'), precision
        ('fp64' or 'mixed': FP32 iterations with FP64 residual replacement),
        delta (replacement threshold of the mixed solve)."""
        self._require_stiffness("solve_static")
        fixed = self.fixed_dofs_array
        u_fixed = self._fixed_dof_values_array(fixed)
        if not np.any(self.F) and not np.any(u_fixed):
            warnings.warn(
                "solve_static: the load vector F is all zeros and no non-zero displacement is "
                "prescribed, so the solution is trivially zero. Add loads first "
                "(add_nodal_force, add_consistent_edge_load, ...).", UserWarning, stacklevel=2)
        x = self.solve_static_device(method=method, **method_kwargs)
        if verbose:
            st = self.last_solve_stats
            print(f"[fea_engine backend=hip] solve: n={self.n_dof - len(self.fixed_dofs)}, nnz=matrix-free, "
                  f"dtype=float64, device={self._op.device}, symmetric=True, spd=n/a, backend=hip, "
                  f"method={st['method']}/{st['precision']}, iterations={st['iterations']}, "
                  f"rel_residual={st['rel_residual']:.3e}, ms={st['elapsed_ms']:.2f}")
        return self.field(x.cpu().numpy(), label="displacement")

    def field(self, data, label=None):
        if not isinstance(data, np.ndarray) or data.ndim not in (1, 2) or data.shape[0] != self.n_dof:
            return data
        return FEField(data, n_nodes=len(self.mesh.nodes), dofs_per_node=3, dof_names=self.dof_names,
                       mesh=self.mesh, label=label)

    # ---- derived results (fe/solver.py:413-430 -> fe/recovery.py) ---------------------------
    _STRESS_NAMES = ("sxx", "syy", "szz", "sxy", "syz", "sxz")
    _STRAIN_NAMES = ("exx", "eyy", "ezz", "gxy", "gyz", "gxz")

    def _recover(self, U, D, kind, at, who):
        """Device recovery (h8_recover_f64): (nodal (N,6), vm_nodal, elem (m,6),
        vm_elem) as numpy; D of the last assemble_stiffness unless given
        (fe/recovery.py:36-42); per-element coefficients become a scale of one D."""
        if self._hex20:
            raise NotImplementedError(f"{who} with Hex20Solid3D is not available on backend='hip'.")
        if at not in ("nodes", "elements"):
            raise ValueError("at must be 'nodes' or 'elements'.")
        U = np.asarray(U)
        if U.ndim != 1 or U.shape[0] != self.n_dof:
            raise ValueError(f"expected a displacement vector of length {self.n_dof}, got shape {U.shape}.")
        if np.iscomplexobj(U):
            raise NotImplementedError("recovery of complex (harmonic) fields: pass U.real or U.imag.")
        Dm = scale = None
        if kind == "stress":
            if D is None:
                args = getattr(self, "_stiffness_args", None)
                if args is None:
                    raise RuntimeError("stress/strain recovery needs the constitutive matrix: call "
                                       "assemble_stiffness() first (it is remembered) or pass D=.")
                D = args[0]
            if has_coefficient(D):
                Dm, scale = factor_scales(D)
                if len(scale) != len(self.mesh.elements):
                    raise ValueError(f"Coefficient.by_element: {len(scale)} values but the mesh has "
                                     f"{len(self.mesh.elements)} elements.")
            else:
                Dm = np.asarray(D, dtype=float)
                if Dm.shape != (6, 6):
                    raise ValueError(f"{who}: D must be 6x6 for Hex8Solid3D, got {Dm.shape}")
        st = self._structure if self._structure is not None else structured_grid(self.mesh)
        if st is None:
            raise NotImplementedError(f"{who}(backend='hip') needs a structured box mesh")
        self._structure = st
        nx, ny, nz, coords0 = st
        from .operator import recover_grid
        out = recover_grid(nx, ny, nz, U, coords0[6] - coords0[0], Dm, scale, nodes=(at == "nodes"),
                           elements=(at == "elements"), device=self.device)
        return tuple(None if t is None else t.cpu().numpy() for t in out)

    def stress(self, U=None, D=None, at="nodes", vectorized=True, backend="numpy", device="cpu"):
        """Stress components [sxx, syy, szz, sxy, syz, sxz]: an FEField of nodal
        values (at="nodes") or an (m, 6) array of element averages
        (fe/solver.py:413-418, fe/recovery.py:378-386); computed on the GPU
        whatever vectorized/backend/device say."""
        nodal, _, elem, _ = self._recover(self.solve_static() if U is None else U, D, "stress", at, "stress")
        if at == "elements":
            return elem
        return FEField(nodal.reshape(-1), n_nodes=len(self.mesh.nodes), dofs_per_node=6,
                       dof_names=self._STRESS_NAMES, mesh=self.mesh, label="stress")

    def strain(self, U=None, at="nodes", vectorized=True, backend="numpy", device="cpu"):
        """Strain components [exx, eyy, ezz, gxy, gyz, gxz] (engineering shear),
        same layout as stress (fe/solver.py:420-423, fe/recovery.py:389-397)."""
        nodal, _, elem, _ = self._recover(self.solve_static() if U is None else U, None, "strain", at, "strain")
        if at == "elements":
            return elem
        return FEField(nodal.reshape(-1), n_nodes=len(self.mesh.nodes), dofs_per_node=6,
                       dof_names=self._STRAIN_NAMES, mesh=self.mesh, label="strain")

    def von_mises(self, U=None, D=None, plane="stress", nu=None, at="nodes", vectorized=True, backend="numpy",
                  device="cpu"):
        """Von Mises stress: a scalar FEField at the nodes or an (m,) array of
        element values (fe/solver.py:425-430, fe/recovery.py:400-419; plane and
        nu matter only for 2-D elements and are validated as there)."""
        if plane not in ("stress", "strain"):
            raise ValueError("plane must be 'stress' or 'strain'.")
        if plane == "strain" and nu is None:
            raise ValueError("von_mises(plane='strain') needs the Poisson ratio nu= to form sigma_zz.")
        _, vm_n, _, vm_e = self._recover(self.solve_static() if U is None else U, D, "stress", at, "von_mises")
        if at == "elements":
            return vm_e
        return FEField(vm_n, n_nodes=len(self.mesh.nodes), dofs_per_node=1, dof_names=("von_mises",),
                       mesh=self.mesh, label="von Mises stress")

    def reactions(self, U=None):
        """Support reactions R = K U - F at the constrained DOFs, zero elsewhere
        (fe/recovery.py:422-448); K U is the device matrix-free product."""
        import torch
        if U is None:
            U = self.solve_static()
        self._require_stiffness("reactions")
        U = np.asarray(U, dtype=float).reshape(-1)
        if U.shape != (self.n_dof,):
            raise ValueError(f"U must have shape ({self.n_dof},), got {U.shape}")
        KU = self._op.matvec(torch.as_tensor(U, device=self._op.device)).cpu().numpy()
        out = np.zeros(self.n_dof)
        fixed = self.fixed_dofs_array
        out[fixed] = KU[fixed] - self.F[fixed]
        return self.field(out, label="reaction")


def fesystem_solve_static_torch(sysobj, method="cg", device=None, tol=1e-8, max_iter=None):
    """fe/torch_sparse_solver.py:268-314: the free-DOF solve of an assembled
    FESystem through the GPU path, returned as a full (n_dof,) numpy array, zero
    at every fixed DOF.  method='cg' is the Jacobi-preconditioned CG of
    fe/torch_sparse_solver.py:165-262 (the same recurrence and stopping rule,
    max_iter 10 * free DOFs); here it runs matrix-free in the HIP kernels.  It
    raises RuntimeError when it does not converge and ValueError on a
    non-positive diagonal, as the reference does.  `device` is accepted for
    signature compatibility: the solve runs on the FESystem's device.  The
    reference solves Ff = F[free] with no Dirichlet correction; a system with
    non-zero prescribed values raises NotImplementedError instead (as
    fe/solver.py:1848-1856 does for its torch backend), and method='dense' (a
    dense torch factorisation) is not part of this backend."""
    if method == "dense":
        raise NotImplementedError("fesystem_solve_static_torch(method='dense') is not supported on the HIP "
                                  "backend; use method='cg'.")
    if method != "cg":
        raise ValueError(f"fesystem_solve_static_torch: unknown method={method!r}, expected 'cg' or 'dense'")
    fixed = sysobj.fixed_dofs_array
    if np.any(sysobj._fixed_dof_values_array(fixed)):
        raise NotImplementedError("fesystem_solve_static_torch: non-zero prescribed displacements are not "
                                  "supported (the reference path drops them); use FESystem.solve_static().")
    x = sysobj.solve_static_device(method="pcg", tol=tol, maxiter=max_iter)
    return x.cpu().numpy()


def _proportional(A, B):
    """c with B = c A (ValueError if not proportional)."""
    piv = np.unravel_index(np.argmax(np.abs(A)), A.shape)
    c = B[piv] / A[piv]
    if np.abs(B - c * A).max() > 1e-12 * max(np.abs(B).max(), 1e-300):
        raise ValueError("accumulate=True needs D proportional to the one already assembled")
    return c


def _materialize_csr(op, elements):
    import scipy.sparse as sp
    import torch
    dev = op.device
    conn = torch.as_tensor(elements, dtype=torch.int64, device=dev)
    g = (3 * conn[:, :, None] + torch.arange(3, device=dev)[None, None, :]).reshape(-1, 24)
    rows = g[:, :, None].expand(-1, 24, 24).reshape(-1)
    cols = g[:, None, :].expand(-1, 24, 24).reshape(-1)
    vals = (op.s_e[:, None, None] * op.K0[None]).reshape(-1)
    K = torch.sparse_coo_tensor(torch.stack([rows, cols]), vals, (op.n_dof, op.n_dof)).coalesce()
    idx = K.indices().cpu().numpy()
    return sp.csr_matrix((K.values().cpu().numpy(), (idx[0], idx[1])), shape=(op.n_dof, op.n_dof))
