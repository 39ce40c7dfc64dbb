#!/usr/bin/env python
"""Benchmark of the H8 linear-elastic PCG path on one MI355X.

Headline workload (the large single solve): 128x32x32 H8 cantilever
(140,481 nodes, 418,176 free DOF), per-element Young's modulus 10**U(9,12)
(seed 2), nu = 0.3, x=0 face clamped, consistent -z traction of 1e6 total on
the x=128 face.  A step is one full solve to a true relative residual < 1e-8
in the chosen form: FP32 PCG iterations preconditioned by the multigrid
V-cycle, with FP64 residual replacement.  The plain FP64 Jacobi-PCG GPU solve
of the same system (the reference's algorithm) is timed beside it.

value = n_free * k64 / t_refined: the FP64 Jacobi-PCG iterations k64 that the
reference's algorithm needs for this system at this tolerance, delivered per
second by the chosen form (DOF x reference iterations per second).  Both forms
solve the same system to the same true residual, so value is proportional to
1 / time-to-solution; the raw PCG throughputs of both forms are reported
beside it.

The target block times the large adjoint workload: 256x64x64 SIMP cantilever
(3,244,800 free DOF), rho ~ U(0.2, 1) (seed 5), p = 3, F_z = -1 at (256,32,0):
the FP32-refined forward solve and the compliance gradient against the plain
FP64 forward solve, with the gradient checked by central differences.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--no-target]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "computation-suite_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "PCG DOF·iters/s + %HBM roofline, FP32-refined vs FP64 @1e-8 rel residual"
HBM_PEAK = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_system(nx, ny, nz, seed):
    import numpy as np
    import h8pcg
    rng = np.random.default_rng(seed)
    m = nx * ny * nz
    E = 10 ** rng.uniform(9, 12, m)
    mesh = h8pcg.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D(), backend="hip", device="cuda:0")
    # per-element E as a scale of one D (E = 1): identical to by_element(D(E_e))
    fs.assemble_stiffness(h8pcg.D_solid3d(h8pcg.Material(E=1.0, nu=0.3)), element_scale=E)
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.add_consistent_facet_load(h8pcg.Hex8Solid3D(), h8pcg.boundary_faces(mesh, 0, float(nx)), 2,
                                 -1e6 / (ny * nz))
    return fs, E


# ---------------------------------------------------------------------------
# CPU legs (the oracle: test infrastructure, used here only as the checker
# and as the timed CPU baseline)
# ---------------------------------------------------------------------------
def cpu_legs(nx, ny, nz, E, F, iters, ref_tol):
    import numpy as np
    import torch
    from threadpoolctl import threadpool_limits
    from oracle import h8_oracle as o
    t0 = time.time()
    nodes, elems = o.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    ke = o.hex8_ke(nodes[elems[0]], o.d_solid3d(1.0, 0.3))
    K = o.assemble_csr(nodes, elems, None, scale=E, ke=np.broadcast_to(ke, (len(elems), 24, 24)))
    clamp = o.nodes_on_plane(nodes, 0, 0.0)
    fixed = [3 * int(n) + c for n in clamp for c in range(3)]
    free, fixed = o.partition(K.shape[0], fixed)
    Kff, Ff = o.reduced_system(K, F, free, fixed)
    log(f"cpu: assembly {time.time() - t0:.1f}s, nnz={Kff.nnz}")
    with threadpool_limits(1):
        M = o.jacobi_preconditioner(Kff)
        t1 = time.perf_counter()
        _, k = o.preconditioned_cg(Kff, Ff, M=M, tol=1e-8, max_iters_hard=iters)
        dt = time.perf_counter() - t1
        base = {"value": len(free) * k / dt, "unit": "DOF·iters/s", "cores": 1, "kind": "port",
                "sample": f"{k} Jacobi-PCG iterations of the same {nx}x{ny}x{nz} system (SciPy CSR, FP64, "
                          f"assembly excluded), {dt:.1f} s", "host_cpu": _cpu_name(), "host_cores": os.cpu_count()}
        # the FP64 reference solution: the oracle's Jacobi-PCG from zero to ref_tol
        t2 = time.perf_counter()
        uf, kref = o.preconditioned_cg(Kff, Ff, M=M, tol=ref_tol, maxiter=20 * len(free))
        log(f"cpu: reference solution, {kref} iterations to {ref_tol:g}, {time.perf_counter() - t2:.1f} s")
    uref = np.zeros(K.shape[0])
    uref[free] = uf
    # the reference's torch backend on the host cores (fe/torch_sparse_solver.py:165-262):
    # CSR FP64 Jacobi-CG, all threads
    nth = os.cpu_count() or 1
    torch.set_num_threads(nth)
    Kt = torch.sparse_csr_tensor(torch.as_tensor(Kff.indptr, dtype=torch.int64),
                                 torch.as_tensor(Kff.indices, dtype=torch.int64),
                                 torch.as_tensor(Kff.data), size=Kff.shape)
    b = torch.as_tensor(Ff)
    dinv = 1.0 / torch.as_tensor(Kff.diagonal())
    x = torch.zeros_like(b)
    r = b.clone()
    z = dinv * r
    p = z.clone()
    rz = torch.dot(r, z)
    t3 = time.perf_counter()
    nt = 30
    for _ in range(nt):
        q = Kt @ p
        a = rz / torch.dot(p, q)
        x += a * p
        r -= a * q
        z = dinv * r
        rzn = torch.dot(r, z)
        p = z + (rzn / rz) * p
        rz = rzn
    dtt = time.perf_counter() - t3
    torch_cpu = {"value": len(free) * nt / dtt, "unit": "DOF·iters/s", "cores": nth, "kind": "port",
                 "sample": f"{nt} Jacobi-CG iterations of the same system, torch CSR FP64 on {nth} threads "
                           f"(the reference's backend='torch', device='cpu' recurrence), {dtt:.2f} s"}
    return K, free, uref, kref, base, torch_cpu


def _cpu_name():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def roofline_block(op, nx, ny, nz, t_iter_s, reps):
    """roofline of the dominant kernel of the chosen form: every kernel of the
    V-cycle timed alone with HIP events on the stream it runs on
    (h8pcg.kbench.time_mg), its share of the measured FP32 iteration = launches
    per cycle x average duration / iteration time; the kernel with the largest
    share is the dominant one.  Also the whole-iteration contract."""
    from h8pcg import kbench
    t = kbench.time_mg(op, reps)
    it_us = 1e6 * t_iter_s
    kern = []
    for slot, k in sorted(t["kernels"].items()):
        if not k["calls"] or not k["avg_us"]:
            continue
        kern.append({"kernel": k["kernel"], "avg_us": k["avg_us"], "bytes": k["bytes"], "calls_per_cycle": k["calls"],
                     "achieved_GBs": k["GBps"], "frac": k["GBps"] * 1e9 / HBM_PEAK,
                     "share_of_iteration": k["calls"] * k["avg_us"] / it_us})
    dom = max(kern, key=lambda k: k["share_of_iteration"])
    it_bytes, _, levels = kbench.mg_iteration_bytes(nx, ny, nz, t["mg_degree"])
    return {"bound": "hbm", "achieved": dom["achieved_GBs"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": dom["frac"], "traffic": None, "kernel": dom["kernel"], "kernel_avg_us": dom["avg_us"],
            "bytes_per_launch": dom["bytes"], "share_of_iteration": dom["share_of_iteration"],
            "timing": f"HIP events around {reps} graph-launched launches of each kernel",
            "other_kernels": [k for k in kern if k is not dom],
            "vcycle_us": t["vcycle_us"], "levels": t["levels"], "level_dofs": [lv[0] for lv in levels],
            "mg_degree": t["mg_degree"],
            "iteration_bytes": it_bytes, "iteration_us": it_us,
            "iteration_frac": it_bytes / t_iter_s / HBM_PEAK}


# ---------------------------------------------------------------------------
# the large adjoint workload (SIMP, 256x64x64)
# ---------------------------------------------------------------------------
def target_block(dims, seed, fd_n, reps):
    import numpy as np
    import torch
    from h8pcg import topopt
    nx, ny, nz = dims
    m = nx * ny * nz
    NJ, NI = ny + 1, nx + 1
    n_dof = 3 * NJ * NI * (nz + 1)
    F = np.zeros(n_dof)
    F[3 * (nx * NJ + ny // 2) + 2] = -1.0
    k, j = np.meshgrid(np.arange(nz + 1), np.arange(NJ), indexing="ij")
    fixed = (3 * (k * NI * NJ + j).ravel()[:, None] + np.arange(3)).ravel()
    n_free = n_dof - len(fixed)
    rng = np.random.default_rng(seed)
    rho_np = rng.uniform(0.2, 1.0, m)
    pb = topopt.SIMPProblem(nx, ny, nz, F, fixed, p=3.0, rho_min=1e-9, device="cuda:0", tol=1e-8,
                            precision="mixed")
    rho = torch.as_tensor(rho_np, device=pb.device)

    def fwd_grad():
        r = rho.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Cv = topopt.compliance(r, pb)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        Cv.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return float(Cv), r.grad.detach(), t1 - t0, t2 - t1

    fwd_grad()                                    # warm-up (graph capture)
    runs = [fwd_grad() for _ in range(reps)]
    C_mixed, g, t_fwd, t_grad = runs[-1]
    t_fwd = min(r[2] for r in runs)
    t_grad = min(r[3] for r in runs)
    st_mixed = dict(pb.last_stats)
    u_mixed = pb.solve(rho, precision="mixed")
    log(f"target: mixed forward {1e3 * t_fwd:.1f} ms, gradient {1e3 * t_grad:.1f} ms, {st_mixed}")
    # plain FP64 Jacobi-PCG forward solve of the same system
    pb.solve(rho, tol=1e-2, precision="fp64")     # graph capture
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    u64 = pb.solve(rho, tol=1e-8, precision="fp64")
    torch.cuda.synchronize()
    t_fp64 = time.perf_counter() - t3
    st64 = dict(pb.last_stats)
    log(f"target: fp64 forward {1e3 * t_fp64:.1f} ms, {st64['iterations']} iterations")
    # FP64 reference solution (GPU FP64 PCG at 1e-12) and the true residual of the mixed one
    uref = pb.solve(rho, tol=1e-12, precision="fp64")
    err = float(torch.linalg.norm(u_mixed - uref) / torch.linalg.norm(uref))
    err64 = float(torch.linalg.norm(u64 - uref) / torch.linalg.norm(uref))
    b = pb.op.masked(pb.F)
    res = float(torch.linalg.norm(pb.op.masked(b, pb.op.matvec(u_mixed))) / torch.linalg.norm(b))
    kernels = roofline_block(pb.op, nx, ny, nz, t_fwd / max(1, st_mixed["inner_iterations"]), 20)
    out = {"workload": f"{nx}x{ny}x{nz} SIMP cantilever, rho~U(0.2,1) (seed {seed}), p=3, rho_min=1e-9, E0=1, "
                       f"nu=0.3, clamped x=0, F_z=-1 at ({nx},{ny // 2},0)",
           "n_free": n_free, "elements": m,
           "forward_ms": 1e3 * t_fwd, "gradient_ms": 1e3 * t_grad, "forward_plus_gradient_ms": 1e3 * (t_fwd + t_grad),
           "fp64_forward_ms": 1e3 * t_fp64, "fp64_iterations": st64["iterations"],
           "speedup_forward_vs_fp64": t_fp64 / t_fwd,
           "speedup_forward_plus_gradient_vs_fp64_forward_plus_gradient": (t_fp64 + t_grad) / (t_fwd + t_grad),
           "fp32_iterations": st_mixed["inner_iterations"], "fp64_replacements": st_mixed["outer"],
           "preconditioner": st_mixed.get("preconditioner"),
           "true_rel_residual": st_mixed["true_rel_residual"], "true_rel_residual_fp64_recomputed": res,
           "rel_err_vs_gpu_fp64_tol1e-12": err, "fp64_rel_err_vs_gpu_fp64_tol1e-12": err64,
           "compliance": C_mixed, "forward_stored_MB": 8 * n_dof / 1e6,
           "roofline": kernels}
    from h8pcg import kbench
    it_b, rep_b, _ = kbench.mg_iteration_bytes(nx, ny, nz, kernels["mg_degree"])
    kernels["solve_bytes"] = st_mixed["inner_iterations"] * it_b + st_mixed["outer"] * rep_b
    kernels["solve_frac"] = kernels["solve_bytes"] / t_fwd / HBM_PEAK
    if fd_n > 0:
        els = np.random.default_rng(seed + 1).choice(m, fd_n, replace=False)
        t4 = time.perf_counter()
        fd, _ = topopt.fd_check(pb, rho, els, h=1e-6, tol=1e-12)
        gg = g.cpu().numpy()[els]
        rel = np.abs(fd - gg) / np.abs(gg)
        out.update({"fd_elements": int(fd_n), "fd_h": 1e-6, "fd_max_rel_diff": float(rel.max()),
                    "fd_median_rel_diff": float(np.median(rel)), "fd_s": time.perf_counter() - t4})
        log(f"target: FD on {fd_n} elements, max rel diff {rel.max():.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-target", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=100)
    ap.add_argument("--ref-tol", type=float, default=1e-12)
    ap.add_argument("--dims", default="128,32,32")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--delta", type=float, default=None,
                    help="residual-replacement threshold (default: the library's per preconditioner)")
    ap.add_argument("--target-dims", default="256,64,64")
    ap.add_argument("--target-seed", type=int, default=5)
    ap.add_argument("--target-fd", type=int, default=32)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {a.gpus} requested); the multi-GPU "
              f"slab decomposition is a design note, not an implemented path.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import h8pcg  # noqa: F401
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    nx, ny, nz = (int(v) for v in a.dims.split(","))
    t0 = time.time()
    fs, E = build_system(nx, ny, nz, a.seed)
    op = fs.operator
    n_free = fs.n_dof - len(fs.fixed_dofs)
    m = nx * ny * nz
    log(f"setup {time.time() - t0:.1f}s: n_dof={fs.n_dof} n_free={n_free} m={m}")
    fs.solve_static_device(tol=1e-2)           # builds dinv (free mask) for the tuner
    tuned = op.autotune(precisions=(64,), cache=False)   # the FP64 Jacobi-PCG kernel choice
    log("autotune", tuned)

    # the load vector resident in HBM before any timed solve (the timed region
    # holds no host transfer)
    F_dev = torch.as_tensor(np.asarray(fs.F, dtype=float), device="cuda:0")

    def solve(precision, tol=1e-8, **kw):
        return fs.solve_static_device(tol=tol, precision=precision, delta=a.delta, F=F_dev, **kw)

    # untimed warm-up: at least --warmup solves, and at least 5 and 0.1 s of them
    # (the refined solve reaches its steady time after about 50 ms of
    # back-to-back solves; the first ones after a pause run up to 10% longer)
    nw, tw = 0, time.perf_counter()
    while nw < max(1, a.warmup) or nw < 5 or time.perf_counter() - tw < 0.1:
        solve("mixed")
        nw += 1
    torch.cuda.synchronize()
    log(f"warmup: {nw} refined solves", fs.last_solve_stats)
    # timed region: K refined solves (multigrid-preconditioned FP32 iterations)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = []
    for _ in range(a.steps):
        u = solve("mixed")
        stats.append(dict(fs.last_solve_stats))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    t_solve = dt / a.steps
    s_ = stats[-1]
    log("refined", s_)

    # plain FP64 Jacobi-PCG GPU path, same system
    solve("fp64")
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    u64 = solve("fp64")
    torch.cuda.synchronize()
    dt64 = time.perf_counter() - t2
    st64 = dict(fs.last_solve_stats)
    log("fp64", st64)
    k64 = st64["iterations"]
    value = n_free * k64 / t_solve

    # the Jacobi form of the refined solve (FP32 iterations without the V-cycle)
    solve("mixed", preconditioner="jacobi")
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    solve("mixed", preconditioner="jacobi")
    torch.cuda.synchronize()
    dtj = time.perf_counter() - t3
    stj = dict(fs.last_solve_stats)

    from h8pcg import kbench
    iters = s_["inner_iterations"]
    roof = roofline_block(op, nx, ny, nz, t_solve / max(1, iters), 50)
    b64 = k64 * (13 * n_free * 8 + 8 * m)
    roof["fp64_iteration_contract_frac"] = b64 / dt64 / HBM_PEAK
    roof["solve_bytes"] = (iters * roof["iteration_bytes"]
                           + s_["outer"] * kbench.mg_iteration_bytes(nx, ny, nz, roof["mg_degree"])[1])
    roof["solve_frac"] = roof["solve_bytes"] / t_solve / HBM_PEAK

    cpu = torch_cpu = None
    acc = {}
    if not a.no_cpu:
        K, free, uref, kref, cpu, torch_cpu = cpu_legs(nx, ny, nz, E, np.asarray(fs.F), a.cpu_iters, a.ref_tol)
        Fn = np.asarray(fs.F)
        for name, uu in (("refined", u), ("fp64", u64)):
            un = uu.cpu().numpy()
            acc[name] = {"rel_err_vs_cpu_fp64": float(np.linalg.norm(un - uref) / np.linalg.norm(uref)),
                         "true_rel_residual_cpu": float(np.linalg.norm((Fn - K @ un)[free]) /
                                                        np.linalg.norm(Fn[free]))}
        acc["cpu_reference"] = f"oracle (SciPy CSR, FP64) Jacobi-PCG from zero to {a.ref_tol:g}, {kref} iterations"

    target = None
    if not a.no_target:
        target = target_block(tuple(int(v) for v in a.target_dims.split(",")), a.target_seed, a.target_fd, 3)

    out = {
        "metric": METRIC, "value": value, "unit": "DOF·iters/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * t_solve, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "value_definition": "n_free x (FP64 Jacobi-PCG iterations of this system to 1e-8, the reference's "
                            "algorithm) / wall time of the FP32-refined solve to the same true residual",
        "config": {"workload": f"{nx}x{ny}x{nz} H8 cantilever, E=10**U(9,12) per element (seed {a.seed}), "
                               f"nu=0.3, clamped x=0, -1e6 consistent tip traction, true rel. residual < 1e-8",
                   "n_free": n_free, "elements": m, "delta": a.delta, "preconditioner": s_.get("preconditioner"),
                   "fp64_kernel": {"variant": op.grid.variant, "layers": op.grid.layers}},
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_threads": torch_cpu,
        "refined": {"fp32_iterations": iters, "fp64_replacements": s_["outer"],
                    "true_rel_residual": s_["true_rel_residual"], "solve_ms": 1e3 * t_solve,
                    "graph_launches": s_["launches"], "pcg_dof_iters_per_s": n_free * iters / t_solve,
                    **acc.get("refined", {})},
        "fp64": {"iterations": k64, "solve_ms": 1e3 * dt64, "dof_iters_per_s": n_free * k64 / dt64,
                 **acc.get("fp64", {})},
        "refined_jacobi": {"fp32_iterations": stj["inner_iterations"], "fp64_replacements": stj["outer"],
                           "solve_ms": 1e3 * dtj},
        "accuracy_reference": acc.get("cpu_reference"),
        "speedup_refined_vs_fp64": dt64 / t_solve,
        "target": target,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
