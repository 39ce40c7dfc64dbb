"""Full-size runs of the SIMP-gradient and batched workloads on one GPU.

    python -m h8pcg.workload_runs simp --dims 96,32,16 --rho 0.3,1.0 --seed 3 --fd 64
    python -m h8pcg.workload_runs simp --dims 256,64,64 --rho 0.2,1.0 --seed 5 --fd 32 --precision mixed
    python -m h8pcg.workload_runs batched --seed 4 --n 256 --tol 1e-8
    python -m h8pcg.workload_runs simploop --dims 96,32,16 --iters 20 --vf 0.4 --radius 1.5
    python -m h8pcg.workload_runs hex20 --dims 64,16,16 --tol 1e-8
    python -m h8pcg.workload_runs refined --dims 128,32,32 --seed 2 --reps 12

simp: SIMP cantilever on an nx x ny x nz unit-cube grid, x = 0 face clamped,
F_z = -1 at the node (nx, ny/2, 0); stiffness scale s = rho_min + (1 - rho_min)
rho^p (p = 3, rho_min = 1e-9, E0 = 1, nu = 0.3) with rho ~ U(lo, hi) per element.
Compliance and dC/drho go through the autograd Function (forward solution
stored); then central differences with h = 1e-6 at `fd` elements drawn by
rng.choice(m, fd, replace=False) (a second generator, seed + 1), each from two
FP64 solves warm-started from u at tol 1e-12, in the cancellation-free form
C(rho + h) - C(rho - h) = -(s+ - s-) u-_e^T K0 u+_e.  Prints one JSON line with
the timings and the largest relative difference between FD and gradient.

hex20: a Hex20 cantilever of nx x ny x nz unit bricks (box_mesh_hex20, E = 210e9,
nu = 0.3, x = 0 clamped, -1e6 in z spread over the x = nx nodes), FP64
Jacobi-PCG to tol; prints DOFs, iterations, device time per iteration and the
average time of the operator K u alone.

refined: the large single solve -- an nx x ny x nz cantilever (workloads.cantilever,
-1e6 consistent tip traction, per-element E = 10**U(9, 12) from seed) -- solved
`reps` times in the chosen form (FP32 iterations preconditioned by the multigrid
V-cycle, FP64 residual replacement, true residual < tol) on a device-resident
load vector; prints the per-solve wall times and the last solve's statistics.
simp --reps n repeats the forward + gradient pass n times and reports the last.

batched: solve_many over batched_cantilevers(seed, n) (nx from {8, 12, .., 48},
ny = nz = nx/4, 4..16 unit tip-traction load cases each, E = 10**U(9, 11) per
element); prints systems, DOF, iterations, device time and DOF*iterations/s.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def _sync():
    import torch
    torch.cuda.synchronize()


def simp(a):
    import torch
    from . import topopt
    nx, ny, nz = (int(v) for v in a.dims.split(","))
    lo, hi = (float(v) for v in a.rho.split(","))
    m = nx * ny * nz
    NJ, NI = ny + 1, nx + 1
    n_dof = 3 * NJ * NI * (nz + 1)
    F = np.zeros(n_dof)
    if ny % 2:
        raise SystemExit("simp: ny must be even (the load node is (nx, ny/2, 0))")
    F[3 * (nx * NJ + ny // 2) + 2] = -1.0
    k, j = np.meshgrid(np.arange(nz + 1), np.arange(NJ), indexing="ij")
    clamp = (k * NI * NJ + j).ravel()
    fixed = (3 * clamp[:, None] + np.arange(3)).ravel()
    rng = np.random.default_rng(a.seed)
    rho_np = rng.uniform(lo, hi, m)
    t0 = time.time()
    pb = topopt.SIMPProblem(nx, ny, nz, F, fixed, p=3.0, rho_min=1e-9, device="cuda:0", tol=a.tol,
                            precision=a.precision)
    pb.op.set_mg_degree(a.mg_degree)
    if a.delta is not None:
        pb.op.DELTA = dict(pb.op.DELTA, mg=a.delta, jacobi=a.delta)
    if a.autotune:
        pb.op.set_free_mask(pb.free_mask)
        pb.op.autotune()
    rho = torch.as_tensor(rho_np, device=pb.device).requires_grad_(True)
    setup_s = time.time() - t0
    for _ in range(max(1, a.reps)):
        rho.grad = None
        _sync()
        t1 = time.perf_counter()
        C = topopt.compliance(rho, pb)
        _sync()
        t2 = time.perf_counter()
        C.backward()
        _sync()
        t3 = time.perf_counter()
    fwd = dict(pb.last_stats)
    g = rho.grad.detach().cpu().numpy()
    out = {"workload": "simp", "dims": [nx, ny, nz], "n_dof": n_dof, "n_free": n_dof - len(fixed), "elements": m,
           "rho": [lo, hi], "seed": a.seed, "precision": a.precision, "tol": a.tol, "setup_s": setup_s,
           "mg_degree": a.mg_degree, "delta": a.delta,
           "reps": max(1, a.reps),
           "compliance": float(C), "forward_ms": 1e3 * (t2 - t1), "gradient_ms": 1e3 * (t3 - t2),
           "forward_stats": {k_: fwd.get(k_) for k_ in ("iterations", "inner_iterations", "outer",
                                                        "rel_residual", "true_rel_residual")}}
    if a.fd > 0:
        els = np.random.default_rng(a.seed + 1).choice(m, a.fd, replace=False)
        t4 = time.perf_counter()
        fd = []
        for c0 in range(0, len(els), 8):          # in chunks, with a progress line per chunk
            part, _ = topopt.fd_check(pb, rho.detach(), els[c0:c0 + 8], h=1e-6, tol=1e-12)
            fd.extend(part)
            print("[simp] fd %d/%d elements, %.1f s" % (c0 + len(part), len(els), time.perf_counter() - t4),
                  file=sys.stderr, flush=True)
        fd = np.array(fd)
        _sync()
        rel = np.abs(fd - g[els]) / np.abs(g[els])
        out.update({"fd_elements": int(a.fd), "fd_h": 1e-6, "fd_s": time.perf_counter() - t4,
                    "fd_max_rel_diff": float(rel.max()), "fd_median_rel_diff": float(np.median(rel))})
    print(json.dumps(out), flush=True)


def refined(a):
    import torch
    from .workloads import cantilever
    nx, ny, nz = (int(v) for v in a.dims.split(","))
    E = 10 ** np.random.default_rng(a.seed).uniform(9, 12, nx * ny * nz)
    fs = cantilever(nx, ny, nz, E_elem=E, device="cuda:0")
    fs.operator.set_mg_degree(a.mg_degree)
    F = torch.as_tensor(np.asarray(fs.F, dtype=float), device="cuda:0")
    times = []
    for _ in range(max(1, a.reps)):
        _sync()
        t = time.perf_counter()
        if a.precision == "mixed":
            fs.solve_static_device(tol=a.tol, precision="mixed", F=F, delta=a.delta)
        else:
            fs.solve_static_device(tol=a.tol, precision="fp64", F=F)
        _sync()
        times.append(1e3 * (time.perf_counter() - t))
    st = fs.last_solve_stats
    out = {"workload": "refined", "dims": [nx, ny, nz], "seed": a.seed, "tol": a.tol, "mg_degree": a.mg_degree,
           "delta": a.delta, "precision": a.precision,
           "n_free": fs.n_dof - len(fs.fixed_dofs), "solve_ms": times,
           "stats": {k: st.get(k) for k in ("iterations", "inner_iterations", "outer", "true_rel_residual",
                                            "elapsed_ms", "launches", "preconditioner")}}
    print(json.dumps(out), flush=True)


def simploop(a):
    """The whole SIMP loop (topopt.topology_optimize_compliance): the simp
    cantilever (x = 0 clamped, F_z = -1 at (nx, ny/2, 0), E0 = 1, nu = 0.3),
    uniform start at the volume fraction, p = 3, rho_min = 1e-3, move 0.2,
    Sigmund filter; per-stage device time."""
    from . import topopt
    from .elements import Hex8Solid3D
    from .material import D_solid3d, Material
    from .mesh import box_mesh
    from .solver import FESystem
    nx, ny, nz = (int(v) for v in a.dims.split(","))
    mesh = box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = FESystem(mesh, Hex8Solid3D(), backend="hip", device="cuda:0")
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    NJ = ny + 1
    F = np.zeros(fs.n_dof)
    F[3 * (nx * NJ + ny // 2) + 2] = -1.0
    D = D_solid3d(Material(E=1.0, nu=0.3))
    tm = {}
    _sync()
    t0 = time.perf_counter()
    res = topopt.topology_optimize_compliance(fs, D, F, a.vf, p=3.0, rho_min=1e-3, move=0.2, n_iter=a.iters,
                                              filter_radius=a.radius, tol=0.0, solve_tol=a.tol,
                                              precision=a.precision, timings=tm)
    _sync()
    wall = time.perf_counter() - t0
    it = len(res["compliance_history"])
    print(json.dumps({"workload": "simploop", "dims": [nx, ny, nz], "elements": nx * ny * nz, "iterations": it,
                      "vf": a.vf, "radius": a.radius, "solve_tol": a.tol, "precision": a.precision,
                      "wall_ms": 1e3 * wall, "stage_ms_per_iteration": {k: 1e3 * v / it for k, v in tm.items()
                                                                        if k not in ("setup", "per_iteration")},
                      "stage_ms_each_iteration": {k: [round(1e3 * x, 3) for x in v]
                                                  for k, v in tm.get("per_iteration", {}).items() if k != "setup"},
                      "compliance_first_last": [res["compliance_history"][0], res["compliance_history"][-1]],
                      "volume": float(res["rho"].mean())}), flush=True)


def batched(a):
    from .batched import solve_many
    from .workloads import batched_cantilevers
    t0 = time.time()
    systems = batched_cantilevers(seed=a.seed, n=a.n, device="cuda:0")
    setup_s = time.time() - t0
    dofs = [fs.n_dof for fs, F in systems for _ in range(F.shape[1])]
    free = [fs.n_dof - len(fs.fixed_dofs) for fs, F in systems for _ in range(F.shape[1])]
    from .batched import FORM_NAMES
    form = None if a.form == "auto" else {v: k for k, v in FORM_NAMES.items()}[a.form]
    solve_many(systems, tol=a.tol, form=form)   # warm-up (the batched autotuner's timing, graph capture)
    _sync()
    t1 = time.perf_counter()
    X, st = solve_many(systems, tol=a.tol, form=form)
    _sync()
    wall = time.perf_counter() - t1
    iters = st["iterations"]
    dof_iters = float(np.dot(free, iters))
    print(json.dumps({"workload": "batched", "seed": a.seed, "instances": a.n, "systems": len(dofs),
                      "total_dof": int(sum(dofs)), "total_free_dof": int(sum(free)), "tol": a.tol,
                      "setup_s": setup_s, "wall_ms": 1e3 * wall, "device_ms": st["elapsed_ms"],
                      "launches": st["launches"], "form": st["form"], "iterations_min": int(iters.min()),
                      "iterations_max": int(iters.max()), "iterations_mean": float(iters.mean()),
                      "all_converged": bool(np.all(st["status"] == 0)),
                      "dof_iters_per_s": dof_iters / wall}), flush=True)


def hex20(a):
    import torch
    from .hex20 import Hex20Solid3D, box_mesh_hex20
    from .material import D_solid3d, Material
    from .solver import FESystem
    nx, ny, nz = (int(v) for v in a.dims.split(","))
    mesh = box_mesh_hex20(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = FESystem(mesh, Hex20Solid3D())
    fs.assemble_stiffness(D_solid3d(Material(E=210e9, nu=0.3)))
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.add_nodal_force(mesh.nodes_on_plane(0, float(nx)), 2, -1e6)
    fs.solve_static_device(tol=a.tol)          # warm-up (graph capture)
    _sync()
    t0 = time.perf_counter()
    fs.solve_static_device(tol=a.tol)
    _sync()
    wall = time.perf_counter() - t0
    st = fs.last_solve_stats
    op = fs.operator
    u = torch.randn(op.n_dof, dtype=torch.float64, device=op.device)
    for _ in range(3):
        op.matvec(u)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.reps):
        op.matvec(u)
    e1.record()
    _sync()
    n_free = fs.n_dof - len(fs.fixed_dofs)
    print(json.dumps({"workload": "hex20", "dims": [nx, ny, nz], "elements": nx * ny * nz, "dof": fs.n_dof,
                      "free_dof": n_free, "tol": a.tol, "iterations": st["iterations"],
                      "rel_residual": st["rel_residual"], "device_ms": st["elapsed_ms"], "wall_ms": 1e3 * wall,
                      "us_per_iteration": 1e3 * st["elapsed_ms"] / max(st["iterations"], 1),
                      "matvec_us": 1e3 * e0.elapsed_time(e1) / a.reps,
                      "dof_iters_per_s": n_free * st["iterations"] / (1e-3 * st["elapsed_ms"])}), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("simp")
    s.add_argument("--dims", default="96,32,16")
    s.add_argument("--rho", default="0.3,1.0")
    s.add_argument("--seed", type=int, default=3)
    s.add_argument("--fd", type=int, default=64)
    s.add_argument("--tol", type=float, default=1e-10)
    s.add_argument("--precision", default="fp64", choices=("fp64", "mixed"))
    s.add_argument("--autotune", action="store_true")
    s.add_argument("--reps", type=int, default=1)
    s.add_argument("--mg-degree", type=int, default=0, help="level-0 Chebyshev steps of the V-cycle (0 = library)")
    s.add_argument("--delta", type=float, default=None, help="residual-replacement threshold of the mixed solve")
    rf = sub.add_parser("refined")
    rf.add_argument("--dims", default="128,32,32")
    rf.add_argument("--seed", type=int, default=2)
    rf.add_argument("--tol", type=float, default=1e-8)
    rf.add_argument("--reps", type=int, default=12)
    rf.add_argument("--mg-degree", type=int, default=0, help="level-0 Chebyshev steps of the V-cycle (0 = library)")
    rf.add_argument("--delta", type=float, default=None, help="residual-replacement threshold (None: the library's)")
    rf.add_argument("--precision", default="mixed", choices=("mixed", "fp64"),
                    help="mixed: the multigrid-preconditioned FP32-refined solve; fp64: Jacobi-PCG in FP64")
    b = sub.add_parser("batched")
    b.add_argument("--seed", type=int, default=4)
    b.add_argument("--n", type=int, default=256)
    b.add_argument("--tol", type=float, default=1e-8)
    b.add_argument("--form", default="auto", choices=("auto", "launched", "persistent-8x16", "persistent-16x16"),
                   help="batched form (auto: the batched autotuner's choice)")
    lp = sub.add_parser("simploop")
    lp.add_argument("--dims", default="96,32,16")
    lp.add_argument("--iters", type=int, default=20)
    lp.add_argument("--vf", type=float, default=0.4)
    lp.add_argument("--radius", type=float, default=1.5)
    lp.add_argument("--tol", type=float, default=1e-8)
    lp.add_argument("--precision", default="fp64", choices=("fp64", "mixed"))
    h = sub.add_parser("hex20")
    h.add_argument("--dims", default="64,16,16")
    h.add_argument("--tol", type=float, default=1e-8)
    h.add_argument("--reps", type=int, default=50)
    a = ap.parse_args(argv)
    {"simp": simp, "batched": batched, "simploop": simploop, "hex20": hex20, "refined": refined}[a.cmd](a)


if __name__ == "__main__":
    main()
