"""GPU parity of the SIMP loop around the gradient: the sensitivity filter
(h8_sens_filter_f64), the OC update (h8_oc_update_f64) and the whole loop
(topopt.topology_optimize_compliance) against the reference's own outputs on
the committed fixture and against the oracle at larger sizes."""
import os
import sys

import numpy as np
import pytest

from oracle import h8_oracle as o

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "topopt_golden.npz")


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


def _t(x):
    import torch
    return torch.as_tensor(np.asarray(x, dtype=float), device="cuda")


def test_filter_matches_reference_and_oracle(h8, gold):
    from h8pcg import topopt
    nx, ny, nz = (int(v) for v in gold["filter/dims"])
    h = gold["filter/h"]
    out = topopt.sensitivity_filter(nx, ny, nz, h, float(gold["filter/radius"]), _t(gold["filter/rho"]),
                                    _t(gold["filter/dc"])).cpu().numpy()
    ref = gold["filter/out"]
    say("filter vs reference", np.abs(out - ref).max() / np.abs(ref).max())
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()
    rng = np.random.default_rng(5)
    # the last radius spans a 13^3 offset box, beyond the kernel's compacted table (direct walk)
    for dims, hh, rad in (((20, 12, 9), (1.0, 0.5, 0.75), 1.7), ((9, 7, 5), (1.0, 1.0, 1.0), 2.5),
                          ((10, 8, 6), (1.0, 1.0, 1.0), 5.5)):
        m = int(np.prod(dims))
        rho, dc = rng.uniform(0.05, 1.0, m), -rng.uniform(0.0, 3.0, m)
        nodes, elems = o.box_mesh(dims[0] * hh[0], dims[1] * hh[1], dims[2] * hh[2], *dims)
        want = o.sensitivity_filter(o.filter_weights(nodes[elems].mean(axis=1), rad), rho, dc)
        got = topopt.sensitivity_filter(*dims, hh, rad, _t(rho), _t(dc)).cpu().numpy()
        say("filter vs oracle", dims, np.abs(got - want).max() / np.abs(want).max())
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_filter_constant_field_full_size(h8):
    import torch
    from h8pcg import topopt
    nx, ny, nz = 128, 32, 32
    m = nx * ny * nz
    out = topopt.sensitivity_filter(nx, ny, nz, (1.0, 1.0, 1.0), 1.5, torch.full((m,), 0.5, device="cuda",
                                    dtype=torch.float64), torch.full((m,), -3.0, device="cuda", dtype=torch.float64))
    assert float((out + 3.0).abs().max()) < 1e-12


def test_oc_update_matches_reference_and_oracle(h8, gold):
    from h8pcg import topopt
    out = topopt.oc_update(_t(gold["oc/rho"]), _t(gold["oc/dc"]), float(gold["oc/vf"]), move=float(gold["oc/move"]),
                           rho_min_bound=float(gold["oc/lo"])).cpu().numpy()
    say("oc vs reference", np.abs(out - gold["oc/out"]).max())
    assert np.abs(out - gold["oc/out"]).max() <= 1e-14
    rng = np.random.default_rng(6)
    m = 1 << 20
    rho = rng.uniform(0.1, 0.9, m)
    dc = -10 ** rng.uniform(-3, 1, m)
    want = o.oc_update(rho, dc, 0.35, move=0.15, rho_min_bound=1e-3)
    got, change = topopt.oc_update(_t(rho), _t(dc), 0.35, move=0.15, rho_min_bound=1e-3, return_change=True)
    got = got.cpu().numpy()
    say("oc vs oracle (1M)", np.abs(got - want).max(), "change", change)
    assert np.abs(got - want).max() <= 1e-14
    assert change == np.abs(got - rho).max()
    assert np.all(got >= rho - 0.15 - 1e-12) and np.all(got <= rho + 0.15 + 1e-12)
    # volume within the reference's own test tolerance (its bisection stops on the multiplier)
    assert np.all(got >= 1e-3) and np.all(got <= 1.0) and abs(got.mean() - 0.35) < 0.02


@pytest.mark.parametrize("sensitivity", ["closed_form", "adjoint"])
def test_topology_run_matches_reference(h8, gold, sensitivity):
    from h8pcg import topopt
    nx, ny, nz = (int(v) for v in gold["run/dims"])
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    assert np.array_equal(fs.fixed_dofs_array, gold["run/fixed"])
    say("topology run", sensitivity)
    res = topopt.topology_optimize_compliance(fs, h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), gold["run/F"], 0.4,
                                              p=3.0, rho_min=1e-2, move=0.2, n_iter=6, filter_radius=1.5,
                                              sensitivity=sensitivity, tol=1e-6, solve_tol=1e-13)
    hist = np.asarray(res["compliance_history"])
    say("hist", hist, "ref", gold["run/hist"])
    assert np.allclose(hist, gold["run/hist"], rtol=1e-8, atol=0)
    assert np.abs(res["rho"] - gold["run/rho"]).max() < 1e-8
    assert np.abs(res["u_final"] - gold["run/u"]).max() <= 1e-8 * np.abs(gold["run/u"]).max()


def test_unknown_sensitivity_rejected(h8):
    from h8pcg import topopt
    mesh = h8.box_mesh(4.0, 2.0, 2.0, 4, 2, 2)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
    with pytest.raises(ValueError, match="unknown sensitivity"):
        topopt.topology_optimize_compliance(fs, np.eye(6), np.zeros(fs.n_dof), 0.4, sensitivity="bogus", n_iter=1)


def test_oc_positive_dc_and_forty_iteration_run(h8):
    """fe/tests/test_topopt.py:139-147 (non-negative sensitivities are clamped,
    not raised on) and 168-185 (40 filtered OC iterations on a short cantilever
    halve the compliance of the uniform start at a held volume), on H8."""
    from h8pcg import topopt
    rho = np.full(5, 0.3)
    dc = np.array([1.0, -1.0, 0.0, -2.0, 1e-15])
    out = topopt.oc_update(_t(rho), _t(dc), 0.3, move=0.2).cpu().numpy()
    assert np.all(np.isfinite(out))
    assert np.abs(out - o.oc_update(rho, dc, 0.3, move=0.2)).max() <= 1e-14

    nx, ny, nz = 10, 5, 2
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    F = np.zeros(fs.n_dof)
    tip = mesh.nodes_on_plane(0, float(nx))
    F[3 * tip + 2] = -1.0 / len(tip)
    say("forty-iteration run")
    res = topopt.topology_optimize_compliance(fs, h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), F, 0.4, p=3.0,
                                              rho_min=1e-2, move=0.2, n_iter=40, filter_radius=1.5,
                                              sensitivity="closed_form")
    hist = np.asarray(res["compliance_history"])
    say("compliance", hist[0], "->", hist[-1], "iterations", len(hist))
    assert len(hist) >= 2 and hist[-1] < 0.5 * hist[0]
    assert abs(res["rho"].mean() - 0.4) < 0.03
    assert np.all(res["rho"] >= 1e-2 - 1e-9) and np.all(res["rho"] <= 1.0 + 1e-9)


def test_mixed_precision_loop_matches_fp64(h8):
    """The SIMP loop with FP32-refined forward solves, warm-started from the
    previous design's displacement like the FP64 loop, follows the FP64 loop's
    compliance history and densities (fe/topopt.py:302-377)."""
    from h8pcg import topopt
    nx, ny, nz = 12, 4, 4
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    out = {}
    for prec in ("fp64", "mixed"):
        fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda")
        fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
        F = np.zeros(fs.n_dof)
        tip = mesh.nodes_on_plane(0, float(nx))
        F[3 * tip + 2] = -1.0 / len(tip)
        say("loop", prec)
        out[prec] = topopt.topology_optimize_compliance(fs, h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), F, 0.4,
                                                        p=3.0, rho_min=1e-2, move=0.2, n_iter=8, filter_radius=1.5,
                                                        sensitivity="closed_form", solve_tol=1e-10,
                                                        precision=prec)
    h64, hm = np.asarray(out["fp64"]["compliance_history"]), np.asarray(out["mixed"]["compliance_history"])
    assert len(h64) == len(hm)
    assert np.allclose(hm, h64, rtol=1e-7, atol=0)
    assert np.abs(out["mixed"]["rho"] - out["fp64"]["rho"]).max() < 1e-6
