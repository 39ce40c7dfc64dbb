"""The 20-node serendipity hexahedron (Hex20Solid3D, fe/elements/solids.py:331-406)
on the HIP backend.

- ``Hex20Solid3D``: the element descriptor (the reference's node order: 8
  corners as Hex8, then the mid-edge nodes 8..19 on the edges (0,1) (1,2) (2,3)
  (3,0) (4,5) (5,6) (6,7) (7,4) (0,4) (1,5) (2,6) (3,7); 3x3x3 Gauss rule);
  ``stiffness_batched`` evaluates sum_gp B^T D B detJ w on the GPU
  (h8_ke20_batched_f64).
- ``box_mesh_hex20``: a structured box of Hex20 bricks with the library's node
  numbering (corners as box_mesh, then the x-, y- and z-edge nodes; see
  csrc/h8_hex20.h).
- ``Hex20Operator``: the matrix-free K u, Jacobi diagonal and FP64 Jacobi-PCG
  of that box (h20_spmv_f64, h20_diag_f64, h20_pcg_f64), which FESystem uses
  for a Hex20 box mesh.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import H8PcgStats, check, ptr
from .mesh import Mesh

# natural coordinates of the 20 nodes (the element's data format)
_XI = np.array([-1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, -1, 1, 1, -1], dtype=float)
_ETA = np.array([-1, -1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, -1, 1, 1], dtype=float)
_ZETA = np.array([-1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, 0, 0, 0, 0], dtype=float)


class Hex20Solid3D:
    n_nodes, dofs_per_node, dim, gauss_order = 20, 3, 3, 3
    # gmsh_conn[:, GMSH_NODE_ORDER] puts a Gmsh Hexahedron20 connectivity into
    # this element's node order (corners agree; mid-edge nodes differ)
    GMSH_NODE_ORDER = [0, 1, 2, 3, 4, 5, 6, 7, 8, 11, 13, 9, 16, 18, 19, 17, 10, 12, 14, 15]

    @staticmethod
    def local_dof_names():
        return ("ux", "uy", "uz")

    def dof_index(self, name):
        names = self.local_dof_names()
        aliases = {"x": 0, "y": 1, "z": 2, "u": 0, "v": 1, "w": 2}
        if isinstance(name, str):
            low = name.lower()
            if low in names:
                return names.index(low)
            if low in aliases:
                return aliases[low]
            raise ValueError(f"unknown DOF name {name!r} for Hex20Solid3D; valid names: {list(names)}")
        return int(name)

    def shape_and_derivs(self, natural_coords):
        """N (20,) and dN/d(xi, eta, zeta) (3, 20) at one natural point (serendipity)."""
        xi, eta, zeta = (float(v) for v in natural_coords)
        N = np.zeros(20)
        dN = np.zeros((3, 20))
        for n in range(20):
            xn, yn, zn = _XI[n], _ETA[n], _ZETA[n]
            if xn and yn and zn:
                a, b, c = xi * xn, eta * yn, zeta * zn
                N[n] = 0.125 * (1 + a) * (1 + b) * (1 + c) * (a + b + c - 2)
                dN[:, n] = (0.125 * xn * (1 + b) * (1 + c) * (2 * a + b + c - 1),
                            0.125 * yn * (1 + a) * (1 + c) * (a + 2 * b + c - 1),
                            0.125 * zn * (1 + a) * (1 + b) * (a + b + 2 * c - 1))
            elif xn == 0:
                N[n] = 0.25 * (1 - xi * xi) * (1 + eta * yn) * (1 + zeta * zn)
                dN[:, n] = (-0.5 * xi * (1 + eta * yn) * (1 + zeta * zn),
                            0.25 * (1 - xi * xi) * yn * (1 + zeta * zn),
                            0.25 * (1 - xi * xi) * (1 + eta * yn) * zn)
            elif yn == 0:
                N[n] = 0.25 * (1 + xi * xn) * (1 - eta * eta) * (1 + zeta * zn)
                dN[:, n] = (0.25 * xn * (1 - eta * eta) * (1 + zeta * zn),
                            -0.5 * eta * (1 + xi * xn) * (1 + zeta * zn),
                            0.25 * (1 + xi * xn) * (1 - eta * eta) * zn)
            else:
                N[n] = 0.25 * (1 + xi * xn) * (1 + eta * yn) * (1 - zeta * zeta)
                dN[:, n] = (0.25 * xn * (1 + eta * yn) * (1 - zeta * zeta),
                            0.25 * (1 + xi * xn) * yn * (1 - zeta * zeta),
                            -0.5 * zeta * (1 + xi * xn) * (1 + eta * yn))
        return N, dN

    def stiffness_batched(self, elem_coords, D, device="cuda"):
        """(m, 60, 60) device tensor for coordinates (m, 20, 3) and D (6, 6) or (m, 6, 6)."""
        import torch
        X = torch.as_tensor(np.asarray(elem_coords, dtype=float), device=device).contiguous()
        if X.ndim != 3 or X.shape[1:] != (20, 3):
            raise ValueError(f"Hex20Solid3D: coordinates must have shape (m, 20, 3), got {tuple(X.shape)}")
        m = X.shape[0]
        Dt = torch.as_tensor(np.asarray(D, dtype=float), device=device).contiguous()
        if Dt.shape == (6, 6):
            stride = 0
        elif Dt.shape == (m, 6, 6):
            stride = 36
        else:
            raise ValueError(f"Hex20Solid3D: D must be (6, 6) or (m, 6, 6), got {tuple(Dt.shape)}")
        ke = torch.empty((m, 60, 60), dtype=torch.float64, device=device)
        with torch.cuda.device(ke.device):
            check(_lib.lib().h8_ke20_batched_f64(ptr(X), ptr(Dt), stride, m, ptr(ke), _lib.stream_handle(ke.device)),
                  "h8_ke20_batched_f64")
        return ke

    def stiffness(self, elem_coords, D, thickness=1.0, gauss_order=None, device="cuda"):
        """One element's 60 x 60 stiffness as a numpy array (computed on the GPU)."""
        if gauss_order not in (None, 3):
            raise NotImplementedError("Hex20Solid3D on the HIP backend integrates with the 3x3x3 rule only")
        ke = self.stiffness_batched(np.asarray(elem_coords, dtype=float)[None], D, device=device)
        return ke[0].cpu().numpy() * thickness


def _hex20_arrays(Lx, Ly, Lz, nx, ny, nz):
    NI, NJ, NK = nx + 1, ny + 1, nz + 1
    nc, nxe, nye = NI * NJ * NK, nx * NJ * NK, NI * ny * NK
    x = np.linspace(0.0, Lx, NI)
    y = np.linspace(0.0, Ly, NJ)
    z = np.linspace(0.0, Lz, NK)
    xm, ym, zm = 0.5 * (x[:-1] + x[1:]), 0.5 * (y[:-1] + y[1:]), 0.5 * (z[:-1] + z[1:])

    def grid(ax, ay, az):       # node coordinates in (k, i, j) order, j fastest
        K, I, J = np.meshgrid(az, ax, ay, indexing="ij")
        return np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)

    nodes = np.vstack([grid(x, y, z), grid(xm, y, z), grid(x, ym, z), grid(x, y, zm)])
    k, i, j = (a.ravel() for a in np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij"))

    def corner(di, dj, dk):
        return ((k + dk) * NI + (i + di)) * NJ + (j + dj)

    def xe(dj, dk):
        return nc + ((k + dk) * nx + i) * NJ + (j + dj)

    def ye(di, dk):
        return nc + nxe + ((k + dk) * NI + (i + di)) * ny + j

    def ze(di, dj):
        return nc + nxe + nye + (k * NI + (i + di)) * NJ + (j + dj)

    cols = [corner(0, 0, 0), corner(1, 0, 0), corner(1, 1, 0), corner(0, 1, 0),
            corner(0, 0, 1), corner(1, 0, 1), corner(1, 1, 1), corner(0, 1, 1),
            xe(0, 0), ye(1, 0), xe(1, 0), ye(0, 0),          # bottom edges (0,1) (1,2) (2,3) (3,0)
            xe(0, 1), ye(1, 1), xe(1, 1), ye(0, 1),          # top edges (4,5) (5,6) (6,7) (7,4)
            ze(0, 0), ze(1, 0), ze(1, 1), ze(0, 1)]          # vertical edges (0,4) (1,5) (2,6) (3,7)
    return nodes, np.stack(cols, axis=1).astype(np.int64)


def box_mesh_hex20(Lx, Ly, Lz, nx, ny, nz):
    """Structured Hex20 box [0,Lx]x[0,Ly]x[0,Lz] with nx x ny x nz bricks: the
    corner nodes numbered as box_mesh (k(nx+1)(ny+1) + i(ny+1) + j), then the
    x-edge, y-edge and z-edge mid-nodes (csrc/h8_hex20.h gives the formulas);
    element e(i,j,k) = k nx ny + i ny + j with the Hex20Solid3D node order."""
    for n in (nx, ny, nz):
        if int(n) != n or n < 1:
            raise ValueError("box_mesh_hex20: nx, ny, nz must be positive integers")
    nodes, elems = _hex20_arrays(float(Lx), float(Ly), float(Lz), int(nx), int(ny), int(nz))
    mesh = Mesh(nodes, elems, dim=3)
    mesh.structure_hex20 = (int(nx), int(ny), int(nz), float(Lx), float(Ly), float(Lz))
    return mesh


def structured_hex20(mesh, rtol=1e-12):
    """(nx, ny, nz, coords of element 0 (20, 3)) if `mesh` has box_mesh_hex20's
    connectivity and congruent bricks, else None."""
    st = getattr(mesh, "structure_hex20", None)
    nodes = np.asarray(mesh.nodes, dtype=float)
    elems = np.asarray(mesh.elements)
    if st is None or elems.ndim != 2 or elems.shape[1] != 20:
        return None
    nx, ny, nz = st[:3]
    ref_nodes, ref_e = _hex20_arrays(1.0, 1.0, 1.0, nx, ny, nz)
    if len(nodes) != len(ref_nodes) or not np.array_equal(elems, ref_e):
        return None
    rel = nodes[elems] - nodes[elems[:, :1]]
    h = rel[0]
    if np.abs(rel - h[None]).max() > rtol * max(np.abs(h).max(), 1e-300) * 16:
        return None
    return nx, ny, nz, nodes[elems[0]]


class Hex20Operator:
    """K = sum_e s_e K0 on a box_mesh_hex20 grid (K0: 60 x 60, the brick's
    stiffness), with the interface FESystem's solve uses (set_free_mask,
    masked, matvec, pcg)."""

    def __init__(self, nx, ny, nz, K0, s_e, device="cuda"):
        import torch
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.device = torch.device(device)
        NI, NJ, NK = self.nx + 1, self.ny + 1, self.nz + 1
        self.n_nodes = NI * NJ * NK + self.nx * NJ * NK + NI * self.ny * NK + NI * NJ * self.nz
        self.n_dof = 3 * self.n_nodes
        self.m = self.nx * self.ny * self.nz
        self.K0buf = torch.as_tensor(K0, dtype=torch.float64, device=self.device).reshape(3600).contiguous()
        self.set_scale(s_e)
        self.dinv = None
        self.free_mask = None
        self._ws = None

    def _stream(self):
        return _lib.stream_handle(self.device)

    def _vec(self, v):
        import torch
        t = v.to(device=self.device, dtype=torch.float64) if isinstance(v, torch.Tensor) else \
            torch.as_tensor(np.asarray(v), dtype=torch.float64, device=self.device)
        t = t.reshape(-1).contiguous()
        if t.numel() != self.n_dof:
            raise ValueError(f"vector has {t.numel()} entries, expected {self.n_dof}")
        return t

    def set_scale(self, s_e):
        import torch
        s = torch.as_tensor(s_e, dtype=torch.float64, device=self.device).reshape(-1)
        if s.numel() != self.m:
            raise ValueError("s_e length mismatch")
        self.s_e = s.contiguous()
        if getattr(self, "free_mask", None) is not None and self.dinv is not None:
            self.set_free_mask(self.free_mask.cpu().numpy())

    def matvec(self, u):
        """K u (every DOF), FP64."""
        import torch
        u = self._vec(u)
        y = torch.empty_like(u)
        with torch.cuda.device(self.device):
            check(_lib.lib().h20_spmv_f64(self.nx, self.ny, self.nz, ptr(self.s_e), ptr(self.K0buf), ptr(u), ptr(y),
                                          self._stream()), "h20_spmv_f64")
        return y

    def set_free_mask(self, free_mask, preconditioner="jacobi", zero_diag_flag=None):
        """dinv: 1/K_ii at free DOFs (jacobi) or 1 (no preconditioner), 0 at fixed."""
        import torch
        fm = torch.as_tensor(np.asarray(free_mask, dtype=np.uint8), device=self.device).reshape(-1).contiguous()
        if fm.numel() != self.n_dof:
            raise ValueError("free mask length mismatch")
        self.free_mask = fm
        dinv = torch.empty(self.n_dof, dtype=torch.float64, device=self.device)
        if preconditioner == "jacobi":
            flag = torch.zeros(1, dtype=torch.int32, device=self.device)
            with torch.cuda.device(self.device):
                check(_lib.lib().h20_diag_f64(self.nx, self.ny, self.nz, ptr(self.s_e), ptr(self.K0buf), ptr(fm),
                                              ptr(dinv), ptr(flag), self._stream()), "")
        elif preconditioner is None or preconditioner == "none":
            dinv.copy_(fm.to(torch.float64))
        else:
            raise ValueError(f"unknown preconditioner {preconditioner!r}")
        self.dinv = dinv
        return dinv

    def masked(self, a, b=None):
        """mask(a - b) with mask = free DOFs."""
        import torch
        out = torch.empty(self.n_dof, dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            check(_lib.lib().h8_masked_sub_f64(ptr(a), ptr(b), ptr(self.dinv), ptr(out), self.n_dof,
                                               self._stream()), "h8_masked_sub_f64")
        return out

    def mg_available(self):
        return False

    def pcg(self, b, tol=1e-8, maxiter=None, x0=None):
        """Jacobi-PCG (FP64, x0 = 0) on the free DOFs; b must already be masked."""
        import torch
        if x0 is not None:
            raise NotImplementedError("Hex20 solves start from x0 = 0 on backend='hip'")
        if self.dinv is None:
            raise RuntimeError("set_free_mask() must be called before solving")
        b = self._vec(b)
        x = torch.empty_like(b)
        if maxiter is None:
            maxiter = 10 * int(self.free_mask.sum().item())
        if self._ws is None:
            nbytes = _lib.lib().h20_pcg_workspace_bytes(self.nx, self.ny, self.nz)
            self._ws = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        st = H8PcgStats()
        with torch.cuda.device(self.device):
            rc = _lib.lib().h20_pcg_f64(self.nx, self.ny, self.nz, ptr(self.s_e), ptr(self.K0buf), ptr(self.dinv),
                                        ptr(b), ptr(x), float(tol), int(maxiter), ptr(self._ws), self._ws.numel(),
                                        C.byref(st), self._stream())
        stats = st.as_dict()
        stats["rc"] = rc
        check(rc)
        return x, stats

    def csr(self, elements):
        """The assembled K (scipy CSR) on the host: COO entries s_e K0 summed."""
        import scipy.sparse as sp
        elems = np.asarray(elements)
        dofs = (3 * elems[:, :, None] + np.arange(3)).reshape(len(elems), 60)
        K0 = self.K0buf.cpu().numpy().reshape(60, 60)
        s = self.s_e.cpu().numpy()
        rows = np.repeat(dofs, 60, axis=1).ravel()
        cols = np.tile(dofs, (1, 60)).ravel()
        vals = (s[:, None, None] * K0[None]).ravel()
        return sp.coo_matrix((vals, (rows, cols)), shape=(self.n_dof, self.n_dof)).tocsr()
