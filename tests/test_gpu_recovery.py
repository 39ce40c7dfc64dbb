"""GPU parity of stress / strain / von Mises recovery (h8_recover_f64 through
FESystem.stress / strain / von_mises) against the reference's own outputs on the
committed fixture, against the oracle on grids of several ragged tiles, and at
128x32x32 through properties that fix the answer (uniform strain, rigid motion)."""
import os
import sys

import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "recovery_golden.npz")


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def h8(gpu):
    import h8pcg
    return h8pcg


def _fesys(h8, dims, lengths, E):
    nx, ny, nz = dims
    mesh = h8.box_mesh(*lengths, nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
    if np.ndim(E) == 0:
        fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=float(E), nu=0.3)))
    else:
        fs.assemble_stiffness(h8.coefficients.by_element([h8.D_solid3d(h8.Material(E=float(e), nu=0.3)) for e in E]))
    return fs


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("name", ["het", "const"])
def test_recovery_matches_reference_fixture(h8, name):
    g = np.load(GOLD)
    dims = tuple(int(v) for v in g[name + "/dims"])
    fs = _fesys(h8, dims, tuple(float(v) for v in g[name + "/lengths"]), g[name + "/E"])
    U = g[name + "/U"]
    say("recovery fixture", name, dims)
    got = {"stress_nodes": np.asarray(fs.stress(U)), "stress_elements": fs.stress(U, at="elements"),
           "strain_nodes": np.asarray(fs.strain(U)), "strain_elements": fs.strain(U, at="elements"),
           "vm_nodes": np.asarray(fs.von_mises(U)), "vm_elements": fs.von_mises(U, at="elements")}
    for k, v in got.items():
        ref = g[f"{name}/{k}"]
        assert v.shape == ref.shape, k
        say(k, _rel(v, ref))
        assert _rel(v, ref) <= 1e-12, k
    S = fs.stress(U)
    assert S.component("sxx").shape == (len(fs.mesh.nodes),)


@pytest.mark.parametrize("dims", [(1, 1, 1), (13, 9, 7), (20, 4, 4), (17, 15, 9)])
def test_recovery_matches_oracle(h8, dims):
    nx, ny, nz = dims
    rng = np.random.default_rng(sum(dims))
    E = 10 ** rng.uniform(9, 12, nx * ny * nz)
    lengths = (nx * 1.0, ny * 0.5, nz * 0.75)
    fs = _fesys(h8, dims, lengths, E)
    U = rng.normal(size=fs.n_dof)
    say("recovery vs oracle", dims)
    nodes, elems = o.box_mesh(*lengths, nx, ny, nz)
    sn, se = o.recover(nodes, elems, U, o.d_solid3d(1.0, 0.3), E)
    en, ee = o.recover(nodes, elems, U, None)
    assert _rel(fs.stress(U).nodal, sn) <= 1e-12
    assert _rel(fs.stress(U, at="elements"), se) <= 1e-12
    assert _rel(fs.strain(U).nodal, en) <= 1e-12
    assert _rel(fs.strain(U, at="elements"), ee) <= 1e-12
    assert _rel(np.asarray(fs.von_mises(U)), o.von_mises(sn)) <= 1e-12
    assert _rel(fs.von_mises(U, at="elements"), o.von_mises(se)) <= 1e-12


def test_recovery_full_size_uniform_strain_and_rigid_motion(h8):
    import torch
    from h8pcg.operator import recover_grid
    nx, ny, nz = 128, 32, 32
    h = (1.0, 0.5, 0.25)
    say("recovery 128x32x32 properties")
    NI, NJ, NK = nx + 1, ny + 1, nz + 1
    k, i, j = np.meshgrid(np.arange(NK), np.arange(NI), np.arange(NJ), indexing="ij")
    X = np.stack([i.ravel() * h[0], j.ravel() * h[1], k.ravel() * h[2]], axis=1)   # node order: y fastest
    rng = np.random.default_rng(9)
    Gm = rng.normal(size=(3, 3))
    U = (X @ Gm.T + rng.normal(size=3)).ravel()
    eps = np.array([Gm[0, 0], Gm[1, 1], Gm[2, 2], Gm[0, 1] + Gm[1, 0], Gm[1, 2] + Gm[2, 1], Gm[0, 2] + Gm[2, 0]])
    D = o.d_solid3d(3.0, 0.25)
    s = 10 ** rng.uniform(0, 2, nx * ny * nz)
    nodal, vm_n, elem, vm_e = (t.cpu().numpy() for t in recover_grid(nx, ny, nz, U, h, D, s, device="cuda"))
    sig = D @ eps
    # elements: s_e D eps exactly; nodes: D eps times the mean scale of the touching elements
    assert _rel(elem, s[:, None] * sig[None, :]) <= 1e-12
    assert _rel(vm_e, s * o.von_mises(sig)) <= 1e-12
    nodal1, _, elem1, _ = (t.cpu().numpy() if t is not None else None
                           for t in recover_grid(nx, ny, nz, U, h, D, None, device="cuda"))
    assert _rel(nodal1, np.broadcast_to(sig, nodal1.shape)) <= 1e-12
    assert _rel(elem1, np.broadcast_to(sig, elem1.shape)) <= 1e-12
    w = rng.normal(size=3)
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    Ur = torch.as_tensor((X @ W.T + rng.normal(size=3)).ravel(), device="cuda")
    nr, _, er, _ = recover_grid(nx, ny, nz, Ur, h, D, s, device="cuda")
    # exact answer 0; rounding of the gradient differences is ~eps |u| / h, far below
    # the O(s |D| |w|) stress a wrong table entry would leave
    tol = 1e-9 * np.abs(s).max() * np.abs(D).max() * float(Ur.abs().max())
    say("rigid", float(nr.abs().max()), float(er.abs().max()), "tol", tol)
    assert float(nr.abs().max()) <= tol and float(er.abs().max()) <= tol


def test_recovery_errors(h8):
    mesh = h8.box_mesh(2.0, 1.0, 1.0, 2, 1, 1)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
    U = np.zeros(fs.n_dof)
    with pytest.raises(RuntimeError):
        fs.stress(U)                                   # no assemble_stiffness, no D=
    assert np.allclose(np.asarray(fs.strain(U)), 0.0)   # strain needs no D
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)))
    with pytest.raises(ValueError):
        fs.stress(U, at="gauss")
    with pytest.raises(ValueError):
        fs.stress(U[:-1])
    with pytest.raises(ValueError):
        fs.von_mises(U, plane="membrane")
