"""Pins of the oracle's multigrid restatement (oracle.mg_levels, mg_hierarchy,
mg_vcycle), the checker the GPU V-cycle is compared with.  Each check is fixed
by mathematics, not by the oracle's own formulas:
  - the prolongation reproduces every linear field exactly (trilinear
    interpolation; a wrong parity, weight or Kronecker order fails);
  - on a free body the Galerkin operators annihilate the six rigid-body modes
    of the coarse grid and are symmetric (P^T A P with the right P);
  - the two Chebyshev steps from zero are the scaled Chebyshev polynomial
    1 - lam p(lam) = T2((theta - lam)/delta) / T2(theta/delta) (Saad, Iterative
    Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1);
  - the V-cycle is a symmetric positive definite operator, and PCG with it
    reaches spsolve's solution in a few iterations;
  - the level plan halves the even axes until at most 128 DOFs.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import h8_oracle as o


def _box(nx, ny, nz, E=None):
    nodes, elems = o.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    ke = o.hex8_ke(nodes[elems[0]], o.d_solid3d(1.0, 0.3))
    scale = np.ones(len(elems)) if E is None else E
    K = o.assemble_csr(nodes, elems, None, scale=scale, ke=np.broadcast_to(ke, (len(elems), 24, 24)))
    return nodes, elems, K.tocsr()


def test_level_plan_halves_even_axes():
    assert [lv[:3] for lv in o.mg_levels(128, 32, 32)] == [(128, 32, 32), (64, 16, 16), (32, 8, 8), (16, 4, 4),
                                                          (8, 2, 2), (4, 1, 1)]
    assert [lv[3:] for lv in o.mg_levels(40, 3, 4)] == [(2, 1, 2), (2, 1, 2), (2, 1, 1), (0, 0, 0)]
    # every level above the coarsest has more than 128 DOFs, the coarsest at most 512
    for dims in [(128, 32, 32), (256, 64, 64), (40, 3, 4), (20, 4, 4)]:
        plan = o.mg_levels(*dims)
        n = [3 * (a + 1) * (b + 1) * (c + 1) for a, b, c, *_ in plan]
        assert all(v > o.MG_COARSE_DOF for v in n[:-1]) and n[-1] <= o.MG_DENSE_MAX
    assert o.mg_levels(7, 5, 3) is None            # nothing to halve: one level


def test_prolongation_reproduces_linear_fields():
    nx, ny, nz = 8, 6, 4
    nodes, _, K = _box(nx, ny, nz)
    free = np.ones(K.shape[0], bool)
    H = o.mg_hierarchy(K, free, (nx, ny, nz))
    P = H[0]["P"]
    cn, _ = o.box_mesh(float(nx), float(ny), float(nz), nx // 2, ny // 2, nz // 2)
    rng = np.random.default_rng(0)
    for _ in range(3):
        a, B = rng.standard_normal(3), rng.standard_normal((3, 3))
        fine = (a + nodes @ B.T).ravel()
        coarse = (a + cn @ B.T).ravel()
        assert np.allclose(P @ coarse, fine, rtol=0, atol=1e-12 * np.abs(fine).max())


def _rigid_modes(nodes):
    n = len(nodes)
    R = np.zeros((3 * n, 6))
    for c in range(3):
        R[c::3, c] = 1.0
    x, y, z = nodes[:, 0], nodes[:, 1], nodes[:, 2]
    R[0::3, 3], R[1::3, 3] = -y, x
    R[1::3, 4], R[2::3, 4] = -z, y
    R[0::3, 5], R[2::3, 5] = z, -x
    return R


def test_galerkin_operators_keep_rigid_modes_and_symmetry():
    nx, ny, nz = 8, 4, 4
    E = 10 ** np.random.default_rng(1).uniform(0, 2, nx * ny * nz)
    nodes, _, K = _box(nx, ny, nz, E)
    free = np.ones(K.shape[0], bool)
    H = o.mg_hierarchy(K, free, (nx, ny, nz))
    dims = (nx, ny, nz)
    for li, lev in enumerate(H):
        A = lev["A"]
        cn, _ = o.box_mesh(float(nx), float(ny), float(nz), *lev["dims"])
        R = _rigid_modes(cn)
        ref = abs(A).max()
        assert abs(A - A.T).max() <= 1e-12 * ref
        assert np.abs(A @ R).max() <= 1e-10 * ref * np.abs(R).max(), li
    assert len(H) >= 3 and dims == H[0]["dims"]


@pytest.mark.parametrize("degree", [2, 3, 4, 5, 6])
def test_chebyshev_steps_are_the_scaled_chebyshev_polynomial(degree):
    nx, ny, nz = 8, 4, 4
    _, _, K = _box(nx, ny, nz)
    free = np.ones(K.shape[0], bool)
    lev = o.mg_hierarchy(K, free, (nx, ny, nz), degree0=degree)[0]
    assert len(lev["cheb"]) == degree
    lam_max = lev["lam"]
    a = lam_max / o.MG_CHEB_RATIO
    theta, delta = 0.5 * (lam_max + a), 0.5 * (lam_max - a)
    lam = np.linspace(0.0, lam_max, 41)
    # a diagonal "matrix" diag(lam) with dinv = 1: `degree` steps from zero on b = 1
    import scipy.sparse as sp
    scal = {"A": sp.diags(lam).tocsr(), "dinv": np.ones_like(lam), "cheb": lev["cheb"]}
    x, _ = o._mg_smooth(scal, np.zeros_like(lam), np.ones_like(lam), np.zeros_like(lam), True)
    # the error factor 1 - lam x(lam) is Tk((theta - lam)/delta) / Tk(theta/delta)
    Tk = np.polynomial.chebyshev.Chebyshev.basis(degree)
    assert np.allclose(1.0 - lam * x, Tk((theta - lam) / delta) / Tk(theta / delta), rtol=0, atol=1e-13)
    # and its largest magnitude on [a, lam_max] is 1 / Tk(theta/delta) (equioscillation)
    on = lam >= a
    assert abs(np.abs(1.0 - lam[on] * x[on]).max() - 1.0 / Tk(theta / delta)) < 1e-2 / Tk(theta / delta)
    # a coarse level keeps two steps whatever the level-0 degree
    coarse = o.mg_hierarchy(K, free, (nx, ny, nz), degree0=degree)[1]
    assert len(coarse["cheb"]) == 2


def test_vcycle_is_spd_and_preconditions_pcg():
    nx, ny, nz = 16, 4, 4
    E = 210e9 * 10 ** np.random.default_rng(2).uniform(-3, 0, nx * ny * nz)
    c = o.cantilever(nx, ny, nz, E_elem=E)
    K, F = c["K"].tocsr(), c["F"]
    free, fixed = o.partition(K.shape[0], c["fixed_dofs"])
    fm = np.zeros(K.shape[0], bool)
    fm[free] = True
    H = o.mg_hierarchy(K, fm, (nx, ny, nz))
    rng = np.random.default_rng(3)
    r1, r2 = rng.standard_normal((2, K.shape[0])) * fm
    for HH in (H, o.mg_hierarchy(K, fm, (nx, ny, nz), degree0=4)):
        z1, z2 = o.mg_vcycle(HH, r1), o.mg_vcycle(HH, r2)
        assert abs(r2 @ z1 - r1 @ z2) <= 1e-11 * np.sqrt((r1 @ z1) * (r2 @ z2))
        assert r1 @ z1 > 0 and r2 @ z2 > 0
        assert np.all(z1[~fm] == 0.0)
    # PCG on the reduced system with M^-1 = V restricted to the free DOFs
    Kff, Ff = o.reduced_system(K, F, free, fixed)
    n = K.shape[0]

    def prec(v):
        full = np.zeros(n)
        full[free] = v
        return o.mg_vcycle(H, full)[free]

    M = spla.LinearOperator(Kff.shape, matvec=prec)
    it = [0]
    uf, info = spla.cg(Kff, Ff, M=M, rtol=1e-10, maxiter=200, callback=lambda _: it.__setitem__(0, it[0] + 1))
    assert info == 0
    ud = spla.spsolve(Kff.tocsc(), Ff)
    assert np.linalg.norm(uf - ud) / np.linalg.norm(ud) < 1e-8
    assert it[0] < 60
