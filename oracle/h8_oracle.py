"""CPU restatement of fea_engine's linear-elastic H8 path, in NumPy/SciPy at FP64.

TEST INFRASTRUCTURE ONLY.  This module is the parity checker for the HIP
library and the CPU baseline that bench.py times beside it.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it, and only
as the checker.  The shipped package (computation-suite_amd/h8pcg) never
imports it and never falls back to it.

Every function names the reference file:line it restates (paths relative to the
reference tree, ``fe/`` = ``fea_engine/src/fea_engine/``).  How each function
is pinned (the checks that would fail on a dropped term, a wrong sign, a wrong
index or a transposed operand) is listed in PINNING below; the tests under
tests/test_oracle_*.py implement those checks.

PINNING (function -> independent checks):
  hex8_shape_and_derivs : partition of unity, Kronecker property at corners,
                          derivative = central FD of N
  hex8_B / hex8_ke      : exact integration by a 3x3x3 rule (2x2x2 is exact for
                          the trilinear brick), ke symmetric, 6 rigid-body
                          null vectors and rank 18, ke proportional to E,
                          uniform-strain energy u^T ke u = V eps^T D eps on a
                          distorted hex, patch test on a distorted mesh
  d_solid3d             : Lame form lambda*(tr)^2 + 2*mu*eps:eps energy identity
  box_mesh              : node/element id formulas, volume = sum detJ
  facet_load_shares     : per-corner share = t*A/4 on a rectangle, total = t*A on a
                          skewed face (closed-form area)
  assemble_csr          : equals a brute-force dense element loop
  dirichlet_partition   : K u = F holds at free DOFs for the full vector
  preconditioned_cg     : converged x equals scipy.sparse.linalg.spsolve; iteration
                          count and residual history equal the reference's own
                          preconditioned_cg on the committed golden fixture
  simp_* / sensitivities: simp_scale_grad vs central FD; closed form vs adjoint
                          solve; both vs cancellation-free FD of the forward solve
  solve_static_batched  : each column equals an independent solve
  eigen_solve           : equals scipy.linalg.pinv(Kff) @ Ff (the minimum-norm
                          solution) on a free and a partly supported H8 body; the
                          count of rigid modes is 6 / 3; K U = F on the full
                          system; an unbalanced load raises LinAlgError
  filter_weights / sensitivity_filter / oc_update / topology_optimize :
                          the reference's own outputs on the committed golden
                          fixture (filter, OC update, a 6-iteration filtered SIMP
                          run); a constant field passes the filter unchanged; the
                          OC result holds the move limit, the bounds and the
                          volume to the bisection tolerance
  hex8_ke_bbar          : the reference's own Hex8SolidBbar stiffness on the committed
                          fixture (unit cube, distorted hex, near-incompressible
                          brick); on an affine displacement field it acts as
                          hex8_ke does (u^T ke_bbar u = V eps^T D eps: B_bar u =
                          B u at every Gauss point); symmetric, the 6 rigid modes
                          in its null space, rank 18
  reduce_constraints / solve_constrained :
                          the reference's own reduced map of an overlapping
                          constraint set and its constrained direct solve on the
                          committed fixture; the solution satisfies every
                          constraint and K u = F at the kept DOFs
  recover / von_mises   : the reference's own recovery outputs on the committed
                          golden fixture (per-element D and one D); a uniform
                          strain field gives D eps at every node and element, a
                          rigid motion gives zero; von Mises of uniaxial stress
                          is |s|, of pure shear sqrt(3)|t|
  element_matvec        : equals assemble_csr(...) @ u for every chunk size
  mg_levels / mg_hierarchy / mg_vcycle (the multigrid preconditioner of the
  FP32 iterations; the reference has none -- its checker role is the GPU
  V-cycle) :              the level plan follows the halving rule; the
                          prolongation reproduces every linear field exactly;
                          on a free body every Galerkin operator annihilates the
                          six rigid-body modes of its grid and is symmetric; k
                          Chebyshev steps (k = 2..6, mg_cheb_coefficients) equal
                          the scaled Chebyshev polynomial
                          Tk((theta - lam)/delta) / Tk(theta/delta) (numpy's
                          Chebyshev basis); the V-cycle is symmetric positive
                          definite (level-0 degree 2 and 4) and PCG with it
                          reaches spsolve's solution (tests/test_oracle_mg.py)
  hex20_shape_and_derivs / hex20_B / hex20_ke / hex20_box_by_midpoints /
  assemble_csr_npe :      partition of unity, Kronecker property at the 20
                          nodes, the isoparametric identity, derivatives = central
                          FD of N; ke symmetric with the 6 rigid-body modes in its
                          null space and rank 54, the constant-strain patch test
                          (energy V eps^T D eps, B u = eps at the Gauss points), the
                          brick integrated exactly by the 3x3x3 rule (equal to
                          4x4x4); the library's box numbering equals the
                          midpoint construction node by node (tests/test_oracle_hex20.py);
                          assemble_csr_npe through the GPU parity of K u
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# ---------------------------------------------------------------------------
# Element: fe/elements/solids.py:130-158, fe/elements/base.py:73-95,268-277,494-514
# ---------------------------------------------------------------------------
# Natural corner coordinates, bottom face counter-clockwise then top face
# (fe/elements/solids.py:135-137).
XI_I = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=float)
ETA_I = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=float)
ZETA_I = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=float)


def gauss_points(n=2):
    """Tensor Gauss-Legendre rule, xi outer / eta middle / zeta inner
    (fe/elements/base.py:73-95)."""
    p, w = np.polynomial.legendre.leggauss(n)
    pts = [(a, b, c) for a in p for b in p for c in p]
    wts = [wa * wb * wc for wa in w for wb in w for wc in w]
    return np.array(pts), np.array(wts)


def hex8_shape_and_derivs(xi, eta, zeta):
    """N (8,) and dN/d(xi,eta,zeta) (3,8) (fe/elements/solids.py:133-142)."""
    N = 0.125 * (1 + xi * XI_I) * (1 + eta * ETA_I) * (1 + zeta * ZETA_I)
    dxi = 0.125 * XI_I * (1 + eta * ETA_I) * (1 + zeta * ZETA_I)
    deta = 0.125 * ETA_I * (1 + xi * XI_I) * (1 + zeta * ZETA_I)
    dzeta = 0.125 * ZETA_I * (1 + xi * XI_I) * (1 + eta * ETA_I)
    return N, np.vstack([dxi, deta, dzeta])


def hex8_B(nat, coords):
    """Strain-displacement matrix (6,24) and signed detJ at one natural point
    (fe/elements/solids.py:144-158; J = dN @ X, fe/elements/base.py:268-277).
    Voigt rows: xx, yy, zz, gamma_xy, gamma_yz, gamma_zx; column 3a+c."""
    _, dN = hex8_shape_and_derivs(*nat)
    J = dN @ coords
    detJ = np.linalg.det(J)
    dNx = np.linalg.solve(J, dN)
    B = np.zeros((6, 24))
    B[0, 0::3] = dNx[0]
    B[1, 1::3] = dNx[1]
    B[2, 2::3] = dNx[2]
    B[3, 0::3] = dNx[1]
    B[3, 1::3] = dNx[0]
    B[4, 1::3] = dNx[2]
    B[4, 2::3] = dNx[1]
    B[5, 0::3] = dNx[2]
    B[5, 2::3] = dNx[0]
    return B, detJ


def hex8_ke(coords, D, order=2):
    """ke = sum_gp B^T D B detJ w (fe/elements/base.py:494-514).  detJ is the
    signed determinant, as in the reference."""
    pts, wts = gauss_points(order)
    ke = np.zeros((24, 24))
    for p, w in zip(pts, wts):
        B, detJ = hex8_B(p, coords)
        ke += (B.T @ D @ B) * detJ * w
    return ke


def hex8_ke_bbar(coords, D):
    """B-bar ke: at every Gauss point B_bar = B - (m m^T/3) B + the element volume
    average of (m m^T/3) B (signed detJ weights), ke = sum B_bar^T D B_bar detJ w
    (fe/elements/solids.py:225-262)."""
    pts, wts = gauss_points(2)
    mm = np.outer([1.0, 1, 1, 0, 0, 0], [1.0, 1, 1, 0, 0, 0]) / 3.0
    Bs, dets = [], []
    Bvol = np.zeros((6, 24))
    V = 0.0
    for p, w in zip(pts, wts):
        B, detJ = hex8_B(p, coords)
        Bs.append(B)
        dets.append(detJ)
        Bvol += mm @ B * detJ * w
        V += detJ * w
    Bvol /= V
    ke = np.zeros((24, 24))
    for B, detJ, w in zip(Bs, dets, wts):
        Bb = B - mm @ B + Bvol
        ke += Bb.T @ D @ Bb * detJ * w
    return ke


def hex8_ke_batched(coords, D):
    """Vectorised ke for many elements: coords (m,8,3), D (6,6) or (m,6,6).
    Same arithmetic as hex8_ke (fe/vectorized_assembly.py:94-214 performs the
    same contraction with one einsum)."""
    coords = np.asarray(coords, dtype=float)
    m = coords.shape[0]
    D = np.asarray(D, dtype=float)
    Dm = np.broadcast_to(D, (m, 6, 6)) if D.ndim == 2 else D
    pts, wts = gauss_points(2)
    ke = np.zeros((m, 24, 24))
    for p, w in zip(pts, wts):
        _, dN = hex8_shape_and_derivs(*p)
        J = np.einsum('an,mnc->mac', dN, coords)          # (m,3,3)
        detJ = np.linalg.det(J)
        dNx = np.linalg.solve(J, np.broadcast_to(dN, (m, 3, 8)))
        B = np.zeros((m, 6, 24))
        B[:, 0, 0::3] = dNx[:, 0]
        B[:, 1, 1::3] = dNx[:, 1]
        B[:, 2, 2::3] = dNx[:, 2]
        B[:, 3, 0::3] = dNx[:, 1]
        B[:, 3, 1::3] = dNx[:, 0]
        B[:, 4, 1::3] = dNx[:, 2]
        B[:, 4, 2::3] = dNx[:, 1]
        B[:, 5, 0::3] = dNx[:, 2]
        B[:, 5, 2::3] = dNx[:, 0]
        ke += np.einsum('msi,mst,mtj->mij', B, Dm, B) * (detJ * w)[:, None, None]
    return ke


# ---------------------------------------------------------------------------
# Material: fe/material.py:25-37,650-659
# ---------------------------------------------------------------------------
def d_solid3d(E, nu):
    """Isotropic 6x6 D, engineering shear on the diagonal (fe/material.py:650-659)."""
    c = E / ((1 + nu) * (1 - 2 * nu))
    g = (1 - 2 * nu) / 2
    return c * np.array([
        [1 - nu, nu, nu, 0, 0, 0],
        [nu, 1 - nu, nu, 0, 0, 0],
        [nu, nu, 1 - nu, 0, 0, 0],
        [0, 0, 0, g, 0, 0],
        [0, 0, 0, 0, g, 0],
        [0, 0, 0, 0, 0, g]])


# ---------------------------------------------------------------------------
# Mesh: fe/mesh.py:716-733 (rectangle), 1270-1285 (extrude), 1288-1289 (box)
# ---------------------------------------------------------------------------
def box_mesh(Lx, Ly, Lz, nx, ny, nz):
    """Nodes (N,3) and elements (m,8) of the structured box.
    node(i,j,k) = k(nx+1)(ny+1) + i(ny+1) + j; x = i*Lx/nx, y = j*Ly/ny,
    z = linspace(0, Lz, nz+1)[k]; element e(i,j,k) = k*nx*ny + i*ny + j with
    corners [n(i,j,k), n(i+1,j,k), n(i+1,j+1,k), n(i,j+1,k)] then the same at k+1."""
    dx, dy = Lx / nx, Ly / ny
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing='ij')
    xy = np.stack([I.ravel() * dx, J.ravel() * dy], axis=1)       # (i major, j minor)
    n2d = (nx + 1) * (ny + 1)
    z = np.linspace(0.0, Lz, nz + 1)
    nodes = np.zeros((n2d * (nz + 1), 3))
    for k in range(nz + 1):
        nodes[k * n2d:(k + 1) * n2d, 0:2] = xy
        nodes[k * n2d:(k + 1) * n2d, 2] = z[k]
    ie, je = np.meshgrid(np.arange(nx), np.arange(ny), indexing='ij')
    ie, je = ie.ravel(), je.ravel()
    nid = lambda i, j: i * (ny + 1) + j
    quad = np.stack([nid(ie, je), nid(ie + 1, je), nid(ie + 1, je + 1), nid(ie, je + 1)], axis=1)
    elems = np.concatenate([np.concatenate([quad + k * n2d, quad + (k + 1) * n2d], axis=1)
                            for k in range(nz)], axis=0)
    return nodes, elems.astype(np.int64)


def nodes_on_plane(nodes, axis, value, tol=1e-9):
    """fe/mesh.py:330-332."""
    return np.where(np.abs(nodes[:, axis] - value) < tol)[0]


# ---------------------------------------------------------------------------
# Facet loads: fe/facet_loads.py:172-179 (Hex8 face table), 283-325 (shares)
# ---------------------------------------------------------------------------
HEX8_FACES = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]


def quad4_face_shares(face_coords, traction):
    """Consistent nodal shares of a uniform scalar traction on one bilinear
    face: sum_gp N * t * sqrt(det(J J^T)) * w over a 2x2 Gauss rule
    (fe/facet_loads.py:283-325; quad4 shape functions at (-1,-1),(1,-1),(1,1),(-1,1))."""
    face_coords = np.asarray(face_coords, dtype=float)
    xa = np.array([-1, 1, 1, -1], dtype=float)
    ea = np.array([-1, -1, 1, 1], dtype=float)
    p, w = np.polynomial.legendre.leggauss(2)
    share = np.zeros(4)
    for a, wa in zip(p, w):
        for b, wb in zip(p, w):
            N = 0.25 * (1 + a * xa) * (1 + b * ea)
            dN = np.vstack([0.25 * xa * (1 + b * ea), 0.25 * ea * (1 + a * xa)])
            J = dN @ face_coords                                   # (2,3)
            meas = np.sqrt(np.linalg.det(J @ J.T))
            share += N * traction * meas * wa * wb
    return share


def add_facet_load(F, nodes, facets, dof_index, traction):
    """FESystem.add_consistent_facet_load for Hex8 faces (fe/solver.py:1229-1270)."""
    for f in facets:
        sh = quad4_face_shares(nodes[list(f)], traction)
        for a, n in enumerate(f):
            F[3 * int(n) + dof_index] += sh[a]
    return F


def tip_face_facets(elements, nodes, x_value, tol=1e-9):
    """All Hex8 local face 3 (nodes 1,2,6,5: the +x face) occurrences whose
    nodes lie on x = x_value, built from element connectivity as the
    reference's docstring shows (fe/solver.py:1246-1252)."""
    face = HEX8_FACES[3]
    out = []
    for conn in elements:
        f = tuple(int(conn[i]) for i in face)
        if np.all(np.abs(nodes[list(f), 0] - x_value) < tol):
            out.append(f)
    return out


def add_nodal_force(F, node_ids, dof_index, total):
    """Even split (fe/solver.py:1178-1206)."""
    share = total / len(node_ids)
    for n in node_ids:
        F[3 * int(n) + dof_index] += share
    return F


# ---------------------------------------------------------------------------
# Assembly: fe/solver.py:683-697 (loop), fe/vectorized_assembly.py:217-261 (COO)
# ---------------------------------------------------------------------------
def element_dof_table(elements):
    """Global DOFs 3*node + c per element, local order 3a+c (fe/solver.py:275-277)."""
    return (3 * elements[:, :, None] + np.arange(3)[None, None, :]).reshape(len(elements), 24)


# ---------------------------------------------------------------------------
# Recovery: fe/recovery.py:234-255 (per-element loop), 258-330 (_recover),
# 364-375 (_mises)
# ---------------------------------------------------------------------------
def recover(nodes, elements, U, D=None, scale=None):
    """(nodal (N,6), element (m,6)) stress sigma = D_e B u (or strain with D=None)
    at the 2x2x2 Gauss points, projected to the corners with the linear shape
    functions: nodal_a = sum N_a val |detJ| w / sum N_a |detJ| w over the elements
    sharing node a; element = volume average (fe/recovery.py:234-255,258-330).
    D_e = scale[e] * D (per-element material as a multiple of one D)."""
    nodes = np.asarray(nodes, dtype=float)
    elements = np.asarray(elements)
    m = len(elements)
    num = np.zeros((len(nodes), 6))
    den = np.zeros(len(nodes))
    elem = np.zeros((m, 6))
    pts, wts = gauss_points(2)
    for i, conn in enumerate(elements):
        coords = nodes[conn]
        ue = U[(3 * conn[:, None] + np.arange(3)[None, :]).ravel()]
        De = None if D is None else (D if scale is None else scale[i] * D)
        acc, vol = np.zeros(6), 0.0
        for p, w in zip(pts, wts):
            B, detJ = hex8_B(p, coords)
            N, _ = hex8_shape_and_derivs(*p)
            val = B @ ue
            if De is not None:
                val = De @ val
            wt = abs(detJ) * w
            num[conn] += wt * N[:, None] * val[None, :]
            den[conn] += wt * N
            acc += wt * val
            vol += wt
        elem[i] = acc / vol
    nodal = np.divide(num, den[:, None], out=np.zeros_like(num), where=den[:, None] > 0)
    return nodal, elem


def von_mises(s):
    """Von Mises stress of (..., 6) Voigt stresses [xx, yy, zz, xy, yz, xz]
    (fe/recovery.py:364-375)."""
    s = np.asarray(s)
    sxx, syy, szz, sxy, syz, sxz = (s[..., j] for j in range(6))
    return np.sqrt(0.5 * ((sxx - syy) ** 2 + (syy - szz) ** 2 + (szz - sxx) ** 2)
                   + 3.0 * (sxy ** 2 + syz ** 2 + sxz ** 2))


def assemble_csr(nodes, elements, D, scale=None, ke=None):
    """Global K as CSR by COO duplicate-sum.  `scale` (m,) multiplies each
    element's ke (per-element E, or the SIMP scale: fe/topopt.py:108-113;
    fe/coefficients.py:148-171 with D_e = E_e/E0 * D)."""
    m = len(elements)
    if ke is None:
        ke = hex8_ke_batched(nodes[elements], D)
    if scale is not None:
        ke = ke * np.asarray(scale, dtype=float)[:, None, None]
    g = element_dof_table(elements)
    rows = np.repeat(g, 24, axis=1).ravel()
    cols = np.tile(g, (1, 24)).ravel()
    n = 3 * len(nodes)
    K = sp.coo_matrix((ke.reshape(m, 576).ravel(), (rows, cols)), shape=(n, n)).tocsr()
    K.sum_duplicates()
    return K


def assemble_dense_loop(nodes, elements, D, scale=None):
    """Brute-force per-element loop into a dense K (fe/solver.py:683-697)."""
    n = 3 * len(nodes)
    K = np.zeros((n, n))
    for e, conn in enumerate(elements):
        ke = hex8_ke(nodes[conn], D)
        if scale is not None:
            ke = ke * scale[e]
        g = np.array([3 * a + c for a in conn for c in range(3)])
        K[np.ix_(g, g)] += ke
    return K


# ---------------------------------------------------------------------------
# Dirichlet partition: fe/solver.py:1399-1401,1481-1509,1977-1979
# ---------------------------------------------------------------------------
def partition(n_dof, fixed_dofs):
    fixed = np.array(sorted(set(int(d) for d in fixed_dofs)), dtype=np.int64)
    mask = np.ones(n_dof, dtype=bool)
    mask[fixed] = False
    free = np.nonzero(mask)[0]
    return free, fixed


def reduced_system(K, F, free, fixed, u_fixed=None):
    """Kff and Ff = F[free] - K[free,fixed] u_fixed (zero shortcut when all 0)."""
    K = sp.csr_matrix(K)
    Kff = K[free][:, free]
    Ff = F[free].astype(float).copy()
    if u_fixed is not None and len(fixed) and np.any(u_fixed):
        Ff = Ff - K[free][:, fixed] @ u_fixed
    return Kff.tocsr(), Ff


# ---------------------------------------------------------------------------
# Jacobi PCG: fe/iterative_solvers.py:193-202 (Jacobi), 362-469 (PCG)
# ---------------------------------------------------------------------------
class NotSPDError(np.linalg.LinAlgError):
    pass


def jacobi_preconditioner(A):
    """fe/iterative_solvers.py:193-202."""
    d = np.asarray(A.diagonal(), dtype=float)
    if np.any(d == 0):
        raise ValueError("jacobi_preconditioner: A has a zero diagonal entry.")
    dinv = 1.0 / d
    return lambda r: dinv * r


def preconditioned_cg(A, b, M=None, x0=None, tol=1e-8, maxiter=None, callback=None,
                      max_iters_hard=None, dtype=np.float64):
    """Hestenes-Stiefel PCG exactly as fe/iterative_solvers.py:434-469:
    x0 = 0, bnorm = ||b|| (1 if 0), early return when ||r||/bnorm < tol, raise
    LinAlgError when p^T A p <= 0, stop when ||r_k||/bnorm < tol, maxiter = 10 n,
    RuntimeError when not converged.  Returns (x, n_iter).

    `max_iters_hard` stops after that many iterations WITHOUT raising and returns
    (x, k) -- used only to time a bounded sample of iterations.  `dtype` lets the
    same recurrence run in float32 to model a reduced-precision inner solve."""
    n = b.shape[0]
    if maxiter is None:
        maxiter = 10 * n
    A_mv = (lambda v: A @ v)
    b = np.asarray(b, dtype=dtype)
    x = np.zeros(n, dtype=dtype) if x0 is None else np.array(x0, dtype=dtype)
    bnorm = np.linalg.norm(b)
    if bnorm == 0.0:
        bnorm = 1.0
    r = b - A_mv(x)
    if np.linalg.norm(r) / bnorm < tol:
        return x, 0
    Minv = M if M is not None else (lambda v: v)
    z = Minv(r)
    p = z.copy()
    rz_old = float(r @ z)
    rel = np.linalg.norm(r) / bnorm
    for k in range(1, maxiter + 1):
        Ap = A_mv(p)
        pAp = float(p @ Ap)
        if pAp <= 0.0:
            raise NotSPDError(f"preconditioned_cg: p^T A p = {pAp:.3e} <= 0 at iteration {k}")
        alpha = rz_old / pAp
        x = x + dtype(alpha) * p
        r = r - dtype(alpha) * Ap
        if callback is not None:
            callback(k, x, r)
        rel = np.linalg.norm(r) / bnorm
        if rel < tol:
            return x, k
        if max_iters_hard is not None and k >= max_iters_hard:
            return x, k
        z = Minv(r)
        rz_new = float(r @ z)
        beta = rz_new / rz_old
        p = z + dtype(beta) * p
        rz_old = rz_new
    raise RuntimeError(f"preconditioned_cg: did not converge in {maxiter} iterations "
                       f"(final relative residual {rel:.3e} >= tol={tol:.1e}).")


def solve_static_pcg(K, F, fixed_dofs, fixed_values=None, tol=1e-8, maxiter=None, callback=None):
    """FESystem._solve_static_iterative with method='pcg', preconditioner='jacobi'
    (fe/solver.py:1922-2029).  Returns (U full, n_iter)."""
    n = K.shape[0]
    free, fixed = partition(n, fixed_dofs)
    u_fixed = None
    if fixed_values is not None:
        u_fixed = np.array([fixed_values.get(int(d), 0.0) for d in fixed])
    Kff, Ff = reduced_system(K, F, free, fixed, u_fixed)
    M = jacobi_preconditioner(Kff)
    uf, it = preconditioned_cg(Kff, Ff, M=M, tol=tol, maxiter=maxiter, callback=callback)
    U = np.zeros(n)
    if u_fixed is not None:
        U[fixed] = u_fixed
    U[free] = uf
    return U, it


def solve_direct(K, F, fixed_dofs, fixed_values=None):
    """Sparse direct solve of the reduced system (the role of
    FESystem._sparse_lu_solve, fe/solver.py:1680-1724)."""
    n = K.shape[0]
    free, fixed = partition(n, fixed_dofs)
    u_fixed = None
    if fixed_values is not None:
        u_fixed = np.array([fixed_values.get(int(d), 0.0) for d in fixed])
    Kff, Ff = reduced_system(K, F, free, fixed, u_fixed)
    U = np.zeros(n)
    if u_fixed is not None:
        U[fixed] = u_fixed
    U[free] = spla.spsolve(Kff.tocsc(), Ff)
    return U


def eigen_solve(Kff, Ff, tol=1e-8):
    """The singular-Kff fallback of the direct solve (fe/solver.py:1583-1678):
    symmetric eigendecomposition; eigenvalues below tol * max(max|lambda|, 1)
    are rigid-body directions.  A load whose projection on them exceeds
    1e-6 max|Ff| raises LinAlgError ("not self-equilibrated"); otherwise those
    directions are dropped, which is the minimum-norm (Moore-Penrose) solution."""
    A = Kff.toarray() if sp.issparse(Kff) else np.asarray(Kff)
    lam, V = np.linalg.eigh(A)
    scale = float(np.max(np.abs(lam))) if lam.size else 1.0
    rigid = np.abs(lam) < tol * max(scale, 1.0)
    if rigid.any():
        res = V[:, rigid].T @ Ff
        mx = float(np.max(np.abs(Ff))) if Ff.size else 0.0
        if mx > 0 and np.max(np.abs(res)) > 1e-6 * mx:
            raise np.linalg.LinAlgError(
                f"singular free-DOF stiffness ({int(rigid.sum())} rigid-body modes) and a load that is not "
                f"self-equilibrated against them: under-constrained model")
    coef = np.where(rigid, 0.0, (V.T @ Ff) / np.where(rigid, 1.0, lam))
    return V @ coef, int(rigid.sum())


# ---------------------------------------------------------------------------
# Batched: fe/batched_solve.py:105-164
# ---------------------------------------------------------------------------
def solve_static_batched(K, F_batch, fixed_dofs, fixed_values=None):
    """One factorisation of Kff, every column with the same Dirichlet correction."""
    n = K.shape[0]
    free, fixed = partition(n, fixed_dofs)
    u_fixed = np.zeros(len(fixed))
    if fixed_values is not None:
        u_fixed = np.array([fixed_values.get(int(d), 0.0) for d in fixed])
    Kff, _ = reduced_system(K, np.zeros(n), free, fixed, None)
    corr = (sp.csr_matrix(K)[free][:, fixed] @ u_fixed) if np.any(u_fixed) else np.zeros(len(free))
    lu = spla.splu(Kff.tocsc())
    X = lu.solve(F_batch[free, :] - corr[:, None])
    U = np.zeros((n, F_batch.shape[1]))
    U[fixed, :] = u_fixed[:, None]
    U[free, :] = X
    return U


# ---------------------------------------------------------------------------
# SIMP: fe/topopt.py:74-90 (scale), 150-155 (compliance), 161-177 (closed form),
# 180-229 + fe/differentiable.py:478-512 (adjoint)
# ---------------------------------------------------------------------------
def simp_scale(rho, p=3.0, rho_min=1e-3):
    rho = np.asarray(rho, dtype=float)
    return rho_min + (1.0 - rho_min) * rho ** p


def simp_scale_grad(rho, p=3.0, rho_min=1e-3):
    rho = np.asarray(rho, dtype=float)
    return p * (1.0 - rho_min) * rho ** (p - 1.0)


def compliance(F, u):
    return float(np.dot(F, u))


def element_quadratic(ke_unit, elements, u, v=None):
    """q_e = u_e^T ke_e v_e (v defaults to u); ke_unit (24,24) or (m,24,24)."""
    g = element_dof_table(elements)
    ue = u[g]
    ve = ue if v is None else v[g]
    if ke_unit.ndim == 2:
        return np.einsum('ei,ij,ej->e', ue, ke_unit, ve)
    return np.einsum('ei,eij,ej->e', ue, ke_unit, ve)


def compliance_sensitivity_closed_form(ke_unit, elements, u, rho, p, rho_min):
    """dC/drho_e = -s'(rho_e) u_e^T ke u_e (fe/topopt.py:161-177)."""
    return -simp_scale_grad(rho, p, rho_min) * element_quadratic(ke_unit, elements, u)


def adjoint_sensitivity(K, ke_unit, elements, fixed_dofs, u, dLdu, rho, p, rho_min):
    """General adjoint (fe/differentiable.py:478-512 applied as fe/topopt.py:180-229):
    K lam = -dL/du on the free DOFs, dL/drho_e = s'(rho_e) lam_e^T ke u_e."""
    n = K.shape[0]
    free, fixed = partition(n, fixed_dofs)
    Kff, _ = reduced_system(K, np.zeros(n), free, fixed)
    lam = np.zeros(n)
    lam[free] = spla.spsolve(Kff.tocsc(), -dLdu[free])
    return simp_scale_grad(rho, p, rho_min) * element_quadratic(ke_unit, elements, lam, u)


# ---------------------------------------------------------------------------
# Workload builders (the configurations timed and checked)
# ---------------------------------------------------------------------------
def cantilever(nx, ny, nz, Lx=None, Ly=None, Lz=None, E=210e9, nu=0.3, load="facet",
               total=-1e6, E_elem=None):
    """Clamped x=0 face, uniform -z traction of `total` on the x=Lx face.
    Returns dict(nodes, elements, K, F, fixed_dofs, D)."""
    Lx = float(nx) if Lx is None else Lx
    Ly = float(ny) if Ly is None else Ly
    Lz = float(nz) if Lz is None else Lz
    nodes, elements = box_mesh(Lx, Ly, Lz, nx, ny, nz)
    D = d_solid3d(1.0 if E_elem is not None else E, nu)
    scale = None if E_elem is None else np.asarray(E_elem, dtype=float)
    K = assemble_csr(nodes, elements, D, scale=scale)
    F = np.zeros(3 * len(nodes))
    if load == "facet":
        facets = tip_face_facets(elements, nodes, Lx)
        add_facet_load(F, nodes, facets, 2, total / (Ly * Lz))
    else:
        add_nodal_force(F, nodes_on_plane(nodes, 0, Lx), 2, total)
    clamp = nodes_on_plane(nodes, 0, 0.0)
    fixed = [3 * int(n) + c for n in clamp for c in range(3)]
    return dict(nodes=nodes, elements=elements, K=K, F=F, fixed_dofs=fixed, D=D)


# ---------------------------------------------------------------------------
# SIMP loop: fe/topopt.py:235-296 (filter, OC update), 302-377 (orchestrator)
# ---------------------------------------------------------------------------
def filter_weights(centroids, radius):
    """Dense H[e, j] = max(0, radius - |c_e - c_j|) (fe/topopt.py:235-243)."""
    diff = centroids[:, None, :] - centroids[None, :, :]
    return np.maximum(0.0, radius - np.sqrt(np.sum(diff ** 2, axis=-1)))


def sensitivity_filter(H, rho, dc, eps=1e-9):
    """(H (rho dc)) / max(max(rho, eps) sum_j H, eps) (fe/topopt.py:246-252)."""
    denom = np.maximum(rho, eps) * H.sum(axis=1)
    return (H @ (rho * dc)) / np.maximum(denom, eps)


def oc_update(rho, dc, volume_fraction, move=0.2, rho_min_bound=1e-3, rho_max_bound=1.0,
              bisection_tol=1e-4, max_bisection_iter=100):
    """Optimality-criteria update with the reference's bracket widening (at
    most 60 doublings of the upper end) and bisection (fe/topopt.py:258-296)."""
    dc = np.minimum(dc, -1e-12)
    l1, l2 = 1e-12, 1.0

    def trial(lam):
        r = np.clip(rho * np.sqrt(-dc / lam), rho - move, rho + move)
        return np.clip(r, rho_min_bound, rho_max_bound)

    for _ in range(60):
        if trial(l2).mean() <= volume_fraction:
            break
        l2 *= 2.0
    rho_new = None
    for _ in range(max_bisection_iter):
        lmid = 0.5 * (l1 + l2)
        rho_new = trial(lmid)
        if rho_new.mean() > volume_fraction:
            l1 = lmid
        else:
            l2 = lmid
        if (l2 - l1) < bisection_tol * max(l2, 1.0):
            break
    return rho_new


def topology_optimize(nodes, elements, D, F, fixed_dofs, volume_fraction, p=3.0, rho_min=1e-3, move=0.2,
                      n_iter=50, filter_radius=None, rho_init=None, tol=1e-3):
    """SIMP compliance minimisation: solve -> closed-form sensitivity ->
    (filter) -> OC update, until max |rho_new - rho| < tol
    (fe/topopt.py:302-377, sensitivity="closed_form"); the forward solve is
    the direct sparse solve.  Returns (rho, compliance history, u)."""
    ke = hex8_ke(nodes[elements[0]], D)
    m = len(elements)
    ke_unit = np.broadcast_to(ke, (m, 24, 24))
    rho = np.full(m, volume_fraction) if rho_init is None else np.array(rho_init, dtype=float)
    H = None
    if filter_radius is not None and filter_radius > 0:
        H = filter_weights(nodes[elements].mean(axis=1), filter_radius)
    hist, u = [], None
    for _ in range(n_iter):
        K = assemble_csr(nodes, elements, None, scale=simp_scale(rho, p, rho_min), ke=ke_unit)
        u = solve_direct(K, F, fixed_dofs)
        hist.append(compliance(F, u))
        dc = compliance_sensitivity_closed_form(ke_unit, elements, u, rho, p, rho_min)
        if H is not None:
            dc = sensitivity_filter(H, rho, dc)
        rho_new = oc_update(rho, dc, volume_fraction, move=move, rho_min_bound=rho_min)
        change = float(np.max(np.abs(rho_new - rho)))
        rho = rho_new
        if change < tol:
            break
    return rho, hist, u


# ---------------------------------------------------------------------------
# Linear constraints: fe/boundary.py:98-166, fe/solver.py:1403-1479
# ---------------------------------------------------------------------------
def reduce_constraints(rows, values, fixed_values, tol=1e-10):
    """(slave, {slave: {master: coef}}, {slave: g}): u_s = g_s - sum coef u_m,
    by Gauss-Jordan elimination with largest-|entry| pivots over all rows at
    once (fe/boundary.py:98-166)."""
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
    cols = sorted({d for r in clean for d in r})
    A = np.zeros((len(clean), len(cols) + 1))
    for i, r in enumerate(clean):
        for d, c in r.items():
            A[i, cols.index(d)] = c
        A[i, -1] = rhs[i]
    scale = max(1.0, np.abs(A).max())
    pivots, used = [], []
    for _ in range(min(len(clean), len(cols))):
        sub = np.abs(A[:, :-1])
        sub[used, :] = 0.0
        for _r, c in pivots:
            sub[:, c] = 0.0
        i, j = np.unravel_index(np.argmax(sub), sub.shape)
        if sub[i, j] <= tol * scale:
            break
        A[i] /= A[i, j]
        for r in range(len(clean)):
            if r != i and A[r, j] != 0.0:
                A[r] -= A[r, j] * A[i]
        pivots.append((i, j))
        used.append(i)
    pc = {j for _, j in pivots}
    slave = [cols[j] for _, j in pivots]
    mat = {cols[j]: {cols[k]: A[i, k] for k in range(len(cols)) if k not in pc and abs(A[i, k]) > 1e-14}
           for i, j in pivots}
    g = {cols[j]: float(A[i, -1]) for i, j in pivots}
    return np.array(slave, dtype=int), mat, g


def solve_constrained(K, F, fixed_dofs, fixed_values, constraints):
    """Direct solve with the constraints eliminated (fe/solver.py:1443-1479):
    u_free = T u_keep + up, T^T K_ff T u_keep = T^T (F_f - K_fc u_c - K_ff up).
    constraints: list of (dofs, coefs, value)."""
    n = K.shape[0]
    fixed = np.array(sorted(set(int(d) for d in fixed_dofs)), dtype=int)
    fv = dict(zip(fixed.tolist(), np.asarray(fixed_values, dtype=float).tolist()))
    free = np.setdiff1d(np.arange(n), fixed)
    rows = [dict(zip(np.asarray(d).tolist(), np.asarray(c, dtype=float).tolist())) for d, c, _ in constraints]
    slave, mat, g = reduce_constraints(rows, [v for _, _, v in constraints], fv)
    keep = np.setdiff1d(free, slave)
    kpos = {int(d): i for i, d in enumerate(keep)}
    T = sp.lil_matrix((n, len(keep)))
    for d in keep:
        T[d, kpos[int(d)]] = 1.0
    for s in slave:
        for m, c in mat[int(s)].items():
            T[int(s), kpos[int(m)]] = -c
    w = np.zeros(n)
    w[fixed] = [fv[int(d)] for d in fixed]
    w[slave] = [g[int(s)] for s in slave]
    T = T.tocsr()
    Kc = (T.T @ K @ T).tocsc()
    x = spla.spsolve(Kc, T.T @ (F - K @ w))
    return T @ x + w


# ---------------------------------------------------------------------------
# Multigrid V-cycle (no counterpart in the reference: the checker of the HIP
# library's multigrid preconditioner, h8_mg_vcycle_f32 / h8_refine_solve_mg).
# The coarse operators here are scipy triple products P^T A P of assembled
# matrices, independent of the library's element-local and stencil forms.
# ---------------------------------------------------------------------------
MG_COARSE_DOF = 128
MG_DENSE_MAX = 512
MG_CHEB_RATIO = 30.0


def _mg_prolong_1d(n, r):
    """(n+1) x (n/r+1) linear interpolation along one axis (r = 1: identity)."""
    if r == 1:
        return sp.identity(n + 1, format="csr")
    rows, cols, vals = [], [], []
    for i in range(n + 1):
        if i % 2 == 0:
            rows.append(i)
            cols.append(i // 2)
            vals.append(1.0)
        else:
            rows += [i, i]
            cols += [i // 2, i // 2 + 1]
            vals += [0.5, 0.5]
    return sp.csr_matrix((vals, (rows, cols)), shape=(n + 1, n // 2 + 1))


def mg_levels(nx, ny, nz):
    """Element counts of the hierarchy: halve the even axes until a level has
    at most MG_COARSE_DOF DOFs or no axis is even.  Returns [(nx, ny, nz, rx,
    ry, rz)] (ratios 0 on the coarsest) or None when the coarsest level has
    more than MG_DENSE_MAX DOFs or there is one level only."""
    out = []
    while True:
        ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        r = [2 if v % 2 == 0 else 1 for v in (nx, ny, nz)]
        if ndof <= MG_COARSE_DOF or r == [1, 1, 1] or len(out) == 11:
            out.append((nx, ny, nz, 0, 0, 0))
            break
        out.append((nx, ny, nz, *r))
        nx, ny, nz = nx // r[0], ny // r[1], nz // r[2]
    last = out[-1]
    if len(out) < 2 or 3 * (last[0] + 1) * (last[1] + 1) * (last[2] + 1) > MG_DENSE_MAX:
        return None
    return out


def mg_hierarchy(K, free, dims, K_abs=None, store32=False, degree0=2):
    """Levels of the V-cycle for the full-length K (n x n) with free mask
    `free` (bool, n) on the box grid dims = (nx, ny, nz) (node order
    k (nx+1)(ny+1) + i (ny+1) + j, DOF 3 node + c).  A_0 = M K M; A_{l+1} =
    P_l^T A_l P_l with P_l = M_l (Pz x Px x Py x I3).  Each level keeps A,
    dinv (0 at masked DOFs), lam (Gershgorin bound of D^-1 A over the free rows,
    times 1.0001; level 0 from K_abs = the assembly of s_e |K0| when given).
    store32: every coarse operator rounded to FP32 as soon as it is formed
    (the library stores them so; the next product and the dense inverse of
    the coarsest level then start from the rounded values).  degree0:
    Chebyshev steps per smoothing pass on level 0 (coarser levels: 2)."""
    plan = mg_levels(*dims)
    if plan is None:
        raise ValueError("grid does not coarsen far enough")
    m = np.asarray(free, dtype=float)
    Mm = sp.diags(m)
    A = (Mm @ K @ Mm).tocsr()
    levels = []
    for li, (nx, ny, nz, rx, ry, rz) in enumerate(plan):
        diag = A.diagonal()
        dinv = np.where(diag > 0, 1.0 / np.where(diag > 0, diag, 1.0), 0.0)
        if li == 0 and K_abs is not None:
            rs = np.asarray(abs(K_abs).sum(axis=1)).ravel()
        else:
            rs = np.asarray(abs(A).sum(axis=1)).ravel()
        lam = 1.0001 * float(np.max(rs * dinv))
        lev = {"A": A, "dinv": dinv, "lam": lam, "dims": (nx, ny, nz)}
        levels.append(lev)
        if rx == 0:
            break
        Pn = sp.kron(sp.kron(_mg_prolong_1d(nz, rz), _mg_prolong_1d(nx, rx)), _mg_prolong_1d(ny, ry))
        P = (sp.diags((dinv != 0).astype(float)) @ sp.kron(Pn, sp.identity(3))).tocsr()
        lev["P"] = P
        A = (P.T @ A @ P).tocsr()
        if store32:
            A.data = A.data.astype(np.float32).astype(np.float64)
    c = levels[-1]
    Ad = c["A"].toarray()
    masked = c["dinv"] == 0
    Ad[masked, :] = 0.0
    Ad[:, masked] = 0.0
    Ad[masked, masked] = 1.0
    c["Ainv"] = np.linalg.inv(Ad)
    for li, lev in enumerate(levels):
        lev["cheb"] = mg_cheb_coefficients(lev["lam"], degree0 if li == 0 else 2)
    return levels


def mg_cheb_coefficients(lam, degree):
    """(ca_k, cb_k) of the Chebyshev steps d = ca d + cb dinv (b - A x), x += d,
    k = 0 .. degree-1, on the interval [lam / MG_CHEB_RATIO, lam] (the
    three-term recurrence of Saad, Iterative Methods for Sparse Linear Systems,
    2nd ed., Alg. 12.1: rho_0 = 1/sigma, rho_k = 1/(2 sigma - rho_{k-1}))."""
    a = lam / MG_CHEB_RATIO
    theta, delta = 0.5 * (lam + a), 0.5 * (lam - a)
    sigma = theta / delta
    out = [(0.0, 1.0 / theta)]
    rho = 1.0 / sigma
    for _ in range(1, degree):
        rn = 1.0 / (2.0 * sigma - rho)
        out.append((rn * rho, 2.0 * rn / delta))
        rho = rn
    return tuple(out)


def _mg_smooth(lev, x, b, d, zero_x):
    """The level's Chebyshev steps d = ca d + cb dinv (b - A x), x += d (Saad,
    Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1)."""
    for s, (ca, cb) in enumerate(lev["cheb"]):
        r = b if (zero_x and s == 0) else b - lev["A"] @ x
        d = (0.0 if (zero_x and s == 0) else ca * d) + cb * lev["dinv"] * r
        x = d.copy() if (zero_x and s == 0) else x + d
    return x, d


def mg_vcycle(levels, r, level=0):
    """z = V(r): Chebyshev pre-smoothing from x = 0, coarse correction with
    the restricted residual, Chebyshev post-smoothing (the same polynomial,
    so the cycle is symmetric)."""
    lev = levels[level]
    if level == len(levels) - 1:
        return lev["Ainv"] @ r
    x, d = _mg_smooth(lev, np.zeros_like(r), r, np.zeros_like(r), True)
    bc = lev["P"].T @ (r - lev["A"] @ x)
    x = x + lev["P"] @ mg_vcycle(levels, bc, level + 1)
    x, _ = _mg_smooth(lev, x, r, d, False)
    return x


def element_matvec(elements, ke, scale, u, chunk=65536):
    """K u = sum_e scale_e ke u_e without forming K (the product the assembly
    of fe/vectorized_assembly.py:217-288 and fe/topopt.py:214-220 defines),
    in chunks of elements; ke one (24, 24) matrix.  Pinned to assemble_csr @ u."""
    u = np.asarray(u, dtype=float)
    out = np.zeros_like(u)
    elements = np.asarray(elements)
    for c0 in range(0, len(elements), chunk):
        el = elements[c0:c0 + chunk]
        dofs = (3 * el[:, :, None] + np.arange(3)).reshape(len(el), 24)
        ye = (u[dofs] @ ke.T) * np.asarray(scale[c0:c0 + chunk], dtype=float)[:, None]
        out += np.bincount(dofs.ravel(), weights=ye.ravel(), minlength=len(u))
    return out


# ---------------------------------------------------------------------------
# Hex20 (20-node serendipity hexahedron): fe/elements/solids.py:331-406 with
# the generic Gauss loop of fe/elements/base.py:494-514 at gauss_order 3
# ---------------------------------------------------------------------------
# natural coordinates of the 20 nodes (fe/elements/solids.py:347-352)
XI20 = np.array([-1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, -1, 1, 1, -1], dtype=float)
ETA20 = np.array([-1, -1, 1, 1, -1, -1, 1, 1, -1, 0, 1, 0, -1, 0, 1, 0, -1, -1, 1, 1], dtype=float)
ZETA20 = np.array([-1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, 0, 0, 0, 0], dtype=float)


def hex20_shape_and_derivs(xi, eta, zeta):
    """N (20,) and dN/d(xi,eta,zeta) (3,20): serendipity shape functions
    (fe/elements/solids.py:354-385), vectorised over the node kinds."""
    a, b, c = xi * XI20, eta * ETA20, zeta * ZETA20
    corner = (XI20 != 0) & (ETA20 != 0) & (ZETA20 != 0)
    ex, ey, ez = XI20 == 0, ETA20 == 0, ZETA20 == 0
    N = np.where(corner, 0.125 * (1 + a) * (1 + b) * (1 + c) * (a + b + c - 2), 0.0)
    N = np.where(ex, 0.25 * (1 - xi ** 2) * (1 + b) * (1 + c), N)
    N = np.where(ey, 0.25 * (1 + a) * (1 - eta ** 2) * (1 + c), N)
    N = np.where(ez, 0.25 * (1 + a) * (1 + b) * (1 - zeta ** 2), N)
    dx = np.where(corner, 0.125 * XI20 * (1 + b) * (1 + c) * (2 * a + b + c - 1), 0.0)
    dy = np.where(corner, 0.125 * ETA20 * (1 + a) * (1 + c) * (a + 2 * b + c - 1), 0.0)
    dz = np.where(corner, 0.125 * ZETA20 * (1 + a) * (1 + b) * (a + b + 2 * c - 1), 0.0)
    dx = np.where(ex, -0.5 * xi * (1 + b) * (1 + c), dx)
    dy = np.where(ex, 0.25 * (1 - xi ** 2) * ETA20 * (1 + c), dy)
    dz = np.where(ex, 0.25 * (1 - xi ** 2) * (1 + b) * ZETA20, dz)
    dx = np.where(ey, 0.25 * XI20 * (1 - eta ** 2) * (1 + c), dx)
    dy = np.where(ey, -0.5 * eta * (1 + a) * (1 + c), dy)
    dz = np.where(ey, 0.25 * (1 + a) * (1 - eta ** 2) * ZETA20, dz)
    dx = np.where(ez, 0.25 * XI20 * (1 + b) * (1 - zeta ** 2), dx)
    dy = np.where(ez, 0.25 * (1 + a) * ETA20 * (1 - zeta ** 2), dy)
    dz = np.where(ez, -0.5 * zeta * (1 + a) * (1 + b), dz)
    return N, np.vstack([dx, dy, dz])


def hex20_B(xi, eta, zeta, X):
    """B (6, 60) and detJ at one point (fe/elements/solids.py:387-401): Voigt
    rows xx, yy, zz, xy, yz, zx, columns 3 a + c."""
    _, dN = hex20_shape_and_derivs(xi, eta, zeta)
    J = dN @ X
    G = np.linalg.solve(J, dN)
    B = np.zeros((6, 60))
    B[0, 0::3], B[1, 1::3], B[2, 2::3] = G[0], G[1], G[2]
    B[3, 0::3], B[3, 1::3] = G[1], G[0]
    B[4, 1::3], B[4, 2::3] = G[2], G[1]
    B[5, 0::3], B[5, 2::3] = G[2], G[0]
    return B, np.linalg.det(J)


def hex20_ke(X, D, order=3):
    """sum over the order^3 Gauss rule of B^T D B detJ w (fe/elements/base.py:494-514)."""
    pts, wts = gauss_points(order)
    ke = np.zeros((60, 60))
    for (xi, eta, zeta), w in zip(pts, wts):
        B, det = hex20_B(xi, eta, zeta, np.asarray(X, dtype=float))
        ke += B.T @ D @ B * det * w
    return ke


def hex20_box_by_midpoints(Lx, Ly, Lz, nx, ny, nz):
    """A Hex20 box built as the reference's own tests build quadratic meshes
    (fe/tests/test_benchmarks_convergence.py:42-62): the Hex8 box, then per
    element the mid-edge nodes at the straight-edge midpoints, numbered in
    order of first appearance and shared through their coordinates."""
    nodes8, elems8 = box_mesh(Lx, Ly, Lz, nx, ny, nz)
    nodes = [tuple(x) for x in nodes8]
    index = {tuple(np.round(x, 9)): i for i, x in enumerate(nodes8)}
    conn = []
    for e in elems8:
        X = nodes8[e]
        row = list(e)
        for a in range(8, 20):
            N8, _ = hex8_shape_and_derivs(XI20[a], ETA20[a], ZETA20[a])
            x = N8 @ X
            key = tuple(np.round(x, 9))
            if key not in index:
                index[key] = len(nodes)
                nodes.append(tuple(x))
            row.append(index[key])
        conn.append(row)
    return np.array(nodes, dtype=float), np.array(conn, dtype=np.int64)


def assemble_csr_npe(n_nodes, elements, ke):
    """Global K (CSR) of elements with any node count: COO duplicate-sum of ke
    (m, 3 npe, 3 npe) over the element DOF tables (3 node + c)."""
    elements = np.asarray(elements)
    m, npe = elements.shape
    nd = 3 * npe
    g = (3 * elements[:, :, None] + np.arange(3)).reshape(m, nd)
    rows = np.repeat(g, nd, axis=1).ravel()
    cols = np.tile(g, (1, nd)).ravel()
    n = 3 * n_nodes
    K = sp.coo_matrix((np.asarray(ke).reshape(m, nd * nd).ravel(), (rows, cols)), shape=(n, n)).tocsr()
    K.sum_duplicates()
    return K
