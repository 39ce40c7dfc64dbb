"""GPU parity of the B-bar Hex8 (Hex8SolidBbar): the element kernel
(h8_ke_bbar_batched_f64) against the reference's own stiffness and the oracle,
and whole solves on the B-bar cantilever of the reference's test against its
direct-solve displacements."""
import os
import sys

import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "bbar_golden.npz")


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def h8(gpu):
    import h8pcg
    return h8pcg


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_bbar_ke_kernel_matches_reference_and_oracle(h8, gold):
    el = h8.Hex8SolidBbar()
    for name in ("cube", "distorted", "brick_incomp"):
        E, nu = gold[f"ke/{name}/E_nu"]
        ke = el.stiffness(gold[f"ke/{name}/X"], o.d_solid3d(E, nu))
        ref = gold[f"ke/{name}/ke"]
        say("bbar ke", name, np.abs(ke - ref).max() / np.abs(ref).max())
        assert np.abs(ke - ref).max() <= 1e-12 * np.abs(ref).max()
    rng = np.random.default_rng(12)
    cube = gold["ke/cube/X"]
    X = cube[None] + rng.uniform(-0.2, 0.2, (37, 8, 3))
    Ds = np.stack([o.d_solid3d(10 ** rng.uniform(0, 3), rng.uniform(0.0, 0.49)) for _ in range(37)])
    ke = el.stiffness_batched(X, Ds).cpu().numpy()
    want = np.stack([o.hex8_ke_bbar(X[i], Ds[i]) for i in range(37)])
    err = np.abs(ke - want).max(axis=(1, 2)) / np.abs(want).max(axis=(1, 2))
    say("bbar ke batch vs oracle", err.max())
    assert err.max() <= 1e-12


@pytest.mark.parametrize("cls,nu", [("Hex8Solid3D", 0.3), ("Hex8Solid3D", 0.45), ("Hex8SolidBbar", 0.3),
                                    ("Hex8SolidBbar", 0.45)])
def test_cantilever_matches_reference_direct_solve(h8, gold, cls, nu):
    key = f"cant/{cls}/{nu}"
    mesh = h8.box_mesh(1.0, 0.05, 0.05, 20, 4, 4)
    el = getattr(h8, cls)()
    fs = h8.FESystem(mesh, el, backend="hip", device="cuda")
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=210e9, nu=nu)))
    tip = mesh.nodes_on_plane(0, 1.0)
    fs.add_nodal_force(tip, 2, -1000.0)
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    ref = gold[key + "/U"]
    U = np.asarray(fs.solve_static(method="pcg", tol=1e-11))
    err = np.linalg.norm(U - ref) / np.linalg.norm(ref)
    tc = int(gold[key + "/tip_node"])
    ratio = U[3 * tc + 2] / (-1000.0 / (3 * 210e9 * 0.05 * 0.05 ** 3 / 12))
    say(key, "iters", fs.last_solve_stats["iterations"], "err", err, "ratio", ratio, float(gold[key + "/ratio"]))
    assert err < 1e-8
    assert abs(ratio - float(gold[key + "/ratio"])) < 1e-8
    Um = np.asarray(fs.solve_static(tol=1e-8, precision="mixed"))
    assert np.linalg.norm(Um - ref) / np.linalg.norm(ref) < 1e-6
