"""Device-resident structured H8 operator and its solves.

A ``GridOperator`` owns, on one GPU: the grid (nx, ny, nz), the reference
element matrix K0 (24x24, computed on the device by the batched element kernel),
the per-element scale s_e, FP32 copies of both for the mixed-precision solve,
and the inverse Jacobi diagonal (zero at fixed DOFs).  Every arithmetic step
runs in libh8pcg.so; torch only allocates device memory and supplies the stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import H8Grid, H8PcgStats, check, ptr


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("h8pcg: no GPU is visible; the HIP backend has no CPU fallback")
    _lib.lib()
    return torch


def device_of(device):
    torch = require_gpu()
    dev = torch.device(device if device is not None else "cuda")
    if dev.type != "cuda":
        raise ValueError(f"h8pcg: device must be a GPU device ('cuda' / 'cuda:N'), got {device!r}")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def element_stiffness_batched(coords, D, device="cuda", bbar=False):
    """ke (m,24,24) for element coordinates (m,8,3) and D (6,6) or (m,6,6),
    computed by the device kernel (bbar: the B-bar form); returns a device tensor."""
    torch = require_gpu()
    dev = device_of(device)
    X = torch.as_tensor(np.ascontiguousarray(coords, dtype=np.float64), device=dev).reshape(-1, 8, 3)
    Dt = torch.as_tensor(np.ascontiguousarray(D, dtype=np.float64), device=dev)
    m = X.shape[0]
    if Dt.dim() == 2:
        if Dt.shape != (6, 6):
            raise ValueError(f"D must be 6x6, got {tuple(Dt.shape)}")
        stride = 0
    else:
        if Dt.shape != (m, 6, 6):
            raise ValueError(f"per-element D must be ({m},6,6), got {tuple(Dt.shape)}")
        stride = 36
    ke = torch.empty((m, 24, 24), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        fn = "h8_ke_bbar_batched_f64" if bbar else "h8_ke_batched_f64"
        check(getattr(_lib.lib(), fn)(ptr(X), ptr(Dt.contiguous()), stride, m, ptr(ke), _lib.stream_handle(dev)),
              fn)
    return ke


_OX = (0, 1, 1, 0, 0, 1, 1, 0)
_OY = (0, 0, 1, 1, 0, 0, 1, 1)
_OZ = (0, 0, 0, 0, 1, 1, 1, 1)


def brick_structured(K0):
    """K0 rebuilt from its first three rows with the brick element's exact
    reflection symmetry, K0[3a+c, 3b+d] = s_c(a) s_d(a) T[a^b][c, d] with
    T[x] = K0[0:3, 3x:3x+3] made to satisfy T[x] = R_x T[x]^T R_x (so the
    result is exactly symmetric).  For an axis-aligned brick this changes K0 by
    rounding only; the node-centric kernel reads just these 72 values."""
    import torch
    K0 = K0.to(torch.float64)
    bits = torch.tensor([_OX, _OY, _OZ], dtype=torch.int64, device=K0.device).T      # (8,3)
    sig = 1.0 - 2.0 * bits.to(torch.float64)                                          # (8,3) reflection signs
    T = K0[0:3].reshape(3, 8, 3).permute(1, 0, 2)                                     # T[x][c][d]
    T = 0.5 * (T + sig[:, :, None] * T.transpose(1, 2) * sig[:, None, :])
    code = bits[:, 0] + 2 * bits[:, 1] + 4 * bits[:, 2]                              # corner -> bit code
    inv = torch.empty(8, dtype=torch.int64, device=K0.device)
    inv[code] = torch.arange(8, device=K0.device)                                     # bit code -> corner
    x = inv[code[:, None] ^ code[None, :]]                                            # (8,8) corner a^b
    blocks = sig[:, None, :, None] * T[x] * sig[:, None, None, :]                    # (a,b,c,d)
    return blocks.permute(0, 2, 1, 3).reshape(24, 24).contiguous()


class GridOperator:
    """K = sum_e s_e K0 on an nx x ny x nz box mesh, matrix-free on one GPU."""

    @property
    def s_e(self):
        """Per-element stiffness scales (FP64 device tensor)."""
        return self._s_e

    @s_e.setter
    def s_e(self, v):
        self._s_e = v
        self._scale_token += 1

    def __init__(self, nx, ny, nz, K0, s_e, device="cuda", variant=_lib.VARIANT_AUTO):
        torch = require_gpu()
        self.device = device_of(device)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        # kernel choice per precision: `grid` for the FP64 operations, `grid32`
        # for the FP32 ones (the refined solve's iterations); AUTO = library default
        self.grid = H8Grid(self.nx, self.ny, self.nz, int(variant), 0)
        self.grid32 = H8Grid(self.nx, self.ny, self.nz, int(variant), 0)
        self.n_nodes = (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.n_dof = 3 * self.n_nodes
        self.m = self.nx * self.ny * self.nz
        K0 = torch.as_tensor(K0, dtype=torch.float64, device=self.device)
        if K0.shape != (24, 24):
            raise ValueError("K0 must be 24x24")
        # library layout: the 24x24 matrix then its 72-entry reflection-sector table
        self.K0buf = torch.empty(_lib.K0_WORDS, dtype=torch.float64, device=self.device)
        self.K0buf[:576] = brick_structured(K0).reshape(-1)
        with torch.cuda.device(self.device):
            check(_lib.lib().h8_sector_table_f64(ptr(self.K0buf), self._stream()), "h8_sector_table_f64")
        self.K0 = self.K0buf[:576].view(24, 24)
        self._scale_token = 0
        self.s_e = torch.as_tensor(s_e, dtype=torch.float64, device=self.device).reshape(-1).contiguous()
        if self.s_e.numel() != self.m:
            raise ValueError(f"s_e has {self.s_e.numel()} entries, the grid has {self.m} elements")
        # FP32 layout: the cast of the FP64 buffer, then the FP64 sector table
        # bit-copied (FP64 element products of FP32 data)
        self.K0_32 = torch.cat([self.K0buf.float(), self.K0buf[576:].contiguous().view(torch.float32)])
        assert self.K0_32.numel() == _lib.K0_WORDS_F32
        self.s_e32 = self.s_e.float().contiguous()
        self.dinv = None
        # set_free_mask's cache: ((preconditioner, scale token), mask, dinv); the
        # token changes whenever s_e is assigned (set_scale or directly: scales
        # are replaced, never edited in place)
        self._mask_cache = None
        self._ws = {}

    # ---- helpers -----------------------------------------------------------
    def _stream(self):
        return _lib.stream_handle(self.device)

    def _vec(self, v, dtype=None):
        import torch
        dtype = dtype or torch.float64
        if isinstance(v, torch.Tensor):
            t = v.to(device=self.device, dtype=dtype)
        else:
            t = torch.as_tensor(np.asarray(v), dtype=dtype, device=self.device)
        t = t.reshape(-1).contiguous()
        if t.numel() != self.n_dof:
            raise ValueError(f"vector has {t.numel()} entries, expected {self.n_dof}")
        return t

    def workspace(self, precision):
        import torch
        ws = self._ws.get(precision)
        if ws is None:
            nbytes = _lib.lib().h8_pcg_workspace_bytes(C.byref(self.grid), precision)
            ws = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
            self._ws[precision] = ws
        return ws

    # ---- operator ------------------------------------------------------------
    def matvec(self, u):
        """K u (every DOF), FP64."""
        import torch
        u = self._vec(u)
        y = torch.empty_like(u)
        with torch.cuda.device(self.device):
            check(_lib.lib().h8_spmv_f64(C.byref(self.grid), ptr(self.s_e), ptr(self.K0buf), ptr(u), ptr(y),
                                         self._stream()), "h8_spmv_f64")
        return y

    def matvec32(self, u):
        """K u in FP32 (FP32 s_e, K0 and vectors)."""
        import torch
        u = self._vec(u, torch.float32)
        y = torch.empty_like(u)
        with torch.cuda.device(self.device):
            check(_lib.lib().h8_spmv_f32(C.byref(self.grid32), ptr(self.s_e32), ptr(self.K0_32), ptr(u), ptr(y),
                                         self._stream()), "h8_spmv_f32")
        return y

    def set_free_mask(self, free_mask, preconditioner="jacobi", zero_diag_flag=None):
        """Builds dinv: 1/K_ii at free DOFs (jacobi) or 1 (no preconditioner), 0 at fixed.
        With zero_diag_flag (a zeroed int32 device tensor shared by many models)
        the zero-diagonal check is left to the caller: the flag becomes 1 and no
        host wait happens here."""
        import torch
        fm_host = np.asarray(free_mask, dtype=np.uint8).reshape(-1)
        if fm_host.size != self.n_dof:
            raise ValueError("free mask length mismatch")
        # the same mask, preconditioner and scales as the last build: dinv is
        # unchanged (no upload, no diagonal kernel, no wait on its zero check)
        key = (preconditioner, self._scale_token)
        c = self._mask_cache
        if (zero_diag_flag is None and c is not None and c[0] == key and c[2] is self.dinv
                and np.array_equal(c[1], fm_host)):
            return self.dinv
        fm = torch.as_tensor(fm_host, device=self.device).contiguous()
        self.free_mask = fm
        # rebuilt in place: solves capture their iterations in graphs keyed on
        # the buffers they read, so a stable address keeps those graphs valid
        dinv = self.dinv if self.dinv is not None and self.dinv.numel() == self.n_dof else \
            torch.empty(self.n_dof, dtype=torch.float64, device=self.device)
        if preconditioner == "jacobi" and zero_diag_flag is not None:
            with torch.cuda.device(self.device):
                check(_lib.lib().h8_diag_async_f64(C.byref(self.grid), ptr(self.s_e), ptr(self.K0buf), ptr(fm),
                                                   ptr(dinv), ptr(zero_diag_flag), self._stream()),
                      "h8_diag_async_f64")
        elif preconditioner == "jacobi":
            flag = torch.zeros(1, dtype=torch.int32, device=self.device)
            with torch.cuda.device(self.device):
                check(_lib.lib().h8_diag_f64(C.byref(self.grid), ptr(self.s_e), ptr(self.K0buf), ptr(fm), ptr(dinv),
                                             ptr(flag), self._stream()), "")
        elif preconditioner is None or preconditioner == "none":
            dinv.copy_(fm.to(torch.float64))
        else:
            raise ValueError(f"unknown preconditioner {preconditioner!r}")
        self.dinv = dinv
        self._mask_cache = (key, fm_host.copy(), dinv) if zero_diag_flag is None else None
        return dinv

    def masked_cols(self, A, b=None):
        """mask(A[c] - b) for every row c of A (ncols, n_dof), one launch."""
        import torch
        A = A.contiguous()
        out = torch.empty_like(A)
        with torch.cuda.device(self.device):
            check(_lib.lib().h8_masked_sub_cols_f64(ptr(A), ptr(b), ptr(self.dinv), ptr(out), self.n_dof,
                                                    A.shape[0], self._stream()), "h8_masked_sub_cols_f64")
        return out

    def masked(self, a, b=None):
        """mask(a - b) with mask = free DOFs."""
        import torch
        out = torch.empty(self.n_dof, dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            check(_lib.lib().h8_masked_sub_f64(ptr(a), ptr(b), ptr(self.dinv), ptr(out), self.n_dof,
                                               self._stream()), "h8_masked_sub_f64")
        return out

    # ---- solves ----------------------------------------------------------------
    def pcg(self, b, tol=1e-8, maxiter=None, x0=None):
        """Jacobi-PCG (FP64) on the free DOFs; b must already be masked.
        Returns (x tensor, stats dict)."""
        import torch
        if self.dinv is None:
            raise RuntimeError("set_free_mask() must be called before solving")
        b = self._vec(b)
        x = torch.zeros_like(b) if x0 is None else self._vec(x0).clone()
        if maxiter is None:
            maxiter = 10 * int(self.free_mask.sum().item())
        ws = self.workspace(64)
        st = H8PcgStats()
        with torch.cuda.device(self.device):
            rc = _lib.lib().h8_pcg_f64(C.byref(self.grid), ptr(self.s_e), ptr(self.K0buf), ptr(self.dinv), ptr(b),
                                       ptr(x), 1 if x0 is not None else 0, float(tol), int(maxiter), ptr(ws),
                                       ws.numel(), C.byref(st), self._stream())
        stats = st.as_dict()
        stats["rc"] = rc
        check(rc)
        return x, stats

    def mg_available(self):
        """True when the grid coarsens far enough for the multigrid preconditioner."""
        return _lib.lib().h8_mg_workspace_bytes(C.byref(self.grid32)) > 0

    def mg_workspace(self):
        import torch
        ws = self._ws.get("mg")
        if ws is None:
            nbytes = _lib.lib().h8_mg_workspace_bytes(C.byref(self.grid32))
            if nbytes == 0:
                raise ValueError(f"multigrid: a {self.nx}x{self.ny}x{self.nz} grid does not coarsen far enough "
                                 f"(halving its even axes)")
            ws = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
            self._ws["mg"] = ws
        return ws

    # residual-replacement threshold of the refined solve per preconditioner:
    # with the V-cycle the FP32 residual falls about 2x per iteration, so 0.1
    # replaces every few iterations (0.5 would replace every other one); with
    # Jacobi, fewer replacements let the FP32 recurrence drift (MI355X sweeps)
    DELTA = {"jacobi": 0.5, "mg": 0.1}

    def refined(self, f, tol=1e-8, maxiter=None, delta=None, x0=None, preconditioner="jacobi"):
        """FP32 PCG iterations with FP64 residual replacement; returns (u, stats).
        x0: initial guess (full length; its fixed-DOF entries are ignored).
        preconditioner: 'jacobi' (z = dinv r) or 'mg' (z = one multigrid
        V-cycle; falls back to 'jacobi' when the coarse operator is singular,
        stats['preconditioner'] says which ran).  delta: replacement threshold
        (None: DELTA[preconditioner])."""
        import torch
        if self.dinv is None:
            raise RuntimeError("set_free_mask() must be called before solving")
        if preconditioner not in ("jacobi", "mg"):
            raise ValueError(f"preconditioner must be 'jacobi' or 'mg', got {preconditioner!r}")
        f = self._vec(f)
        u = torch.zeros_like(f) if x0 is None else self._vec(x0).clone()
        if maxiter is None:
            maxiter = 10 * int(self.free_mask.sum().item())
        ws = self.workspace(32)
        st = H8PcgStats()
        used = preconditioner
        d_mg = float(self.DELTA["mg"] if delta is None else delta)
        d_j = float(self.DELTA["jacobi"] if delta is None else delta)
        with torch.cuda.device(self.device):
            rc = None
            if preconditioner == "mg":
                mws = self.mg_workspace()
                rc = _lib.lib().h8_refine_solve_mg(C.byref(self.grid32), ptr(self.s_e), ptr(self.K0buf),
                                                   ptr(self.s_e32), ptr(self.K0_32), ptr(self.dinv), ptr(f), ptr(u),
                                                   float(tol), int(maxiter), d_mg, 0 if x0 is None else 1,
                                                   ptr(ws), ws.numel(), ptr(mws), mws.numel(), C.byref(st),
                                                   self._stream())
                if rc == _lib.H8_MG_DECLINED:
                    used = "jacobi"
                    if x0 is not None:
                        u = self._vec(x0).clone()
            if used == "jacobi":
                fn = _lib.lib().h8_refine_solve if x0 is None else _lib.lib().h8_refine_solve_x0
                rc = fn(C.byref(self.grid32), ptr(self.s_e), ptr(self.K0buf), ptr(self.s_e32), ptr(self.K0_32),
                        ptr(self.dinv), ptr(f), ptr(u), float(tol), int(maxiter), d_j, ptr(ws), ws.numel(),
                        C.byref(st), self._stream())
        stats = st.as_dict()
        stats["rc"] = rc
        stats["preconditioner"] = used
        check(rc)
        return u, stats

    def mg_vcycle(self, r):
        """z = V(r): one V-cycle of the multigrid preconditioner (FP32), after
        building the hierarchy for the current scales and free mask."""
        import torch
        if self.dinv is None:
            raise RuntimeError("set_free_mask() must be called first")
        r = self._vec(r, torch.float32)
        z = torch.zeros_like(r)
        ws = self.workspace(32)
        mws = self.mg_workspace()
        with torch.cuda.device(self.device):
            check(_lib.lib().h8_mg_vcycle_f32(C.byref(self.grid32), ptr(self.s_e), ptr(self.K0buf), ptr(self.s_e32),
                                              ptr(self.K0_32), ptr(self.dinv), ptr(r), ptr(z), ptr(ws), ws.numel(),
                                              ptr(mws), mws.numel(), self._stream()), "h8_mg_vcycle_f32")
        return z

    def element_bilinear(self, u, v, w):
        """out_e = w_e u_e^T K0 v_e (device)."""
        import torch
        u = self._vec(u)
        v = self._vec(v)
        w = torch.as_tensor(w, dtype=torch.float64, device=self.device).reshape(-1).contiguous()
        if w.numel() != self.m:
            raise ValueError("w must have one entry per element")
        out = torch.empty(self.m, dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            check(_lib.lib().h8_elem_bilinear_f64(C.byref(self.grid), ptr(self.K0buf), ptr(u), ptr(v), ptr(w),
                                                  ptr(out), self._stream()), "h8_elem_bilinear_f64")
        return out

    def set_variant(self, variant, layers=0, precision=None):
        """Kernel variant (and sector-kernel layers per block) for FP64 (64), FP32
        (32) or both (None) operations."""
        for p, g in ((64, self.grid), (32, self.grid32)):
            if precision in (None, p):
                g.variant = int(variant)
                g.layers = int(layers)

    def recover(self, u, spacing, D=None, scale=None, nodes=True, elements=True):
        """Stress / strain recovery on this grid (see recover_grid)."""
        return recover_grid(self.nx, self.ny, self.nz, u, spacing, D, scale, nodes, elements, self.device)

    def set_arith(self, arith):
        """Element-product arithmetic of the FP32 sector and resident kernels:
        0 (library choice: FP32), 32 (FP32 products) or 64 (FP64 products of
        the FP32 data, outputs rounded to FP32)."""
        if int(arith) not in (0, 32, 64):
            raise ValueError(f"arith must be 0, 32 or 64, got {arith}")
        self.grid32.arith = int(arith)

    def set_mg_degree(self, degree):
        """Chebyshev steps per smoothing pass on the finest level of the
        multigrid preconditioner: 0 (library choice: 4 up to 2^20 DOFs, 3
        above) or 2..6.  Coarser levels always take 2 steps."""
        if int(degree) != 0 and not 2 <= int(degree) <= 6:
            raise ValueError(f"mg_degree must be 0 (library choice) or 2..6, got {degree}")
        self.grid32.mg_degree = int(degree)

    def autotune(self, precisions=(32, 64), cache=True):
        """Times the candidate kernels for this grid and keeps the fastest per
        precision (see h8pcg.autotune).  Returns {precision: (variant, layers, us)}."""
        from .autotune import tune
        out = {}
        for p in precisions:
            v, l, us = tune(self, p, cache=cache)
            self.set_variant(v, l, p)
            out[p] = (v, l, us)
        return out

    def update_scale_(self, s):
        """Element scales replaced in place (the FP64 and FP32 buffers keep their
        addresses, so captured solve graphs stay valid); the Jacobi diagonal is
        marked stale and rebuilt by the next set_free_mask."""
        import torch
        s = torch.as_tensor(s, dtype=torch.float64, device=self.device).reshape(-1)
        if s.numel() != self.m:
            raise ValueError("s_e length mismatch")
        self._s_e.copy_(s)
        self.s_e32.copy_(s)
        self._scale_token += 1

    def set_scale(self, s_e):
        self.update_scale_(s_e)
        if getattr(self, "free_mask", None) is not None and self.dinv is not None:
            self.set_free_mask(self.free_mask.cpu().numpy())


def recover_grid(nx, ny, nz, u, spacing, D=None, scale=None, nodes=True, elements=True, device="cuda"):
    """Stress (D: 6x6) or strain (D None) recovery on the device for the box grid
    nx x ny x nz with brick spacing (hx, hy, hz): the lumped projection of the
    2x2x2 Gauss-point values to the nodes and the element volume averages, with
    their von Mises values (fe/recovery.py:234-420).  scale: per-element multiple
    of D (None = 1).  Returns device tensors (nodal (N,6), vm_nodal (N,),
    elem (m,6), vm_elem (m,)), None where not asked for."""
    import torch
    hx, hy, hz = (float(v) for v in spacing)
    grid = H8Grid(int(nx), int(ny), int(nz), 0, 0)
    n_nodes, m = (nx + 1) * (ny + 1) * (nz + 1), nx * ny * nz
    f64 = dict(dtype=torch.float64, device=device)
    uu = (u.to(**f64) if isinstance(u, torch.Tensor) else torch.as_tensor(np.asarray(u), **f64)).reshape(-1)
    uu = uu.contiguous()
    if uu.numel() != 3 * n_nodes:
        raise ValueError(f"vector has {uu.numel()} entries, expected {3 * n_nodes}")
    Dt = None
    if D is not None:
        Dt = (D.to(**f64) if isinstance(D, torch.Tensor) else torch.as_tensor(np.asarray(D, dtype=float), **f64))
        if tuple(Dt.shape) != (6, 6):
            raise ValueError(f"D must be 6x6 for Hex8Solid3D, got {tuple(Dt.shape)}")
        Dt = Dt.contiguous()
    st = None
    if scale is not None:
        st = (scale.to(**f64) if isinstance(scale, torch.Tensor)
              else torch.as_tensor(np.asarray(scale, dtype=float), **f64)).reshape(-1).contiguous()
        if st.numel() != m:
            raise ValueError(f"scale has {st.numel()} entries, the grid has {m} elements")
    tab = torch.empty(_lib.RECOVERY_WORDS, **f64)
    nodal = torch.empty((n_nodes, 6), **f64) if nodes else None
    vm_n = torch.empty(n_nodes, **f64) if nodes else None
    elem = torch.empty((m, 6), **f64) if elements else None
    vm_e = torch.empty(m, **f64) if elements else None
    with torch.cuda.device(device):
        stream = _lib.stream_handle(device)
        check(_lib.lib().h8_recovery_table_f64(hx, hy, hz, ptr(Dt) or None, ptr(tab), stream),
              "h8_recovery_table_f64")
        check(_lib.lib().h8_recover_f64(C.byref(grid), ptr(tab), ptr(st) or None, ptr(uu), ptr(nodal) or None,
                                        ptr(vm_n) or None, ptr(elem) or None, ptr(vm_e) or None, stream),
              "h8_recover_f64")
    return nodal, vm_n, elem, vm_e
