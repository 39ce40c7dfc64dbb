"""Kernel micro-benchmark for profiling the two PCG iteration kernels.

    python -m h8pcg.kbench --dims 256,64,64 --reps 50 [--precision 32,64] [--variant V] [--layers L] [--copy-mib 256]
    python -m h8pcg.kbench --dims 256,64,64 --reps 50 --recovery
    python -m h8pcg.kbench --dims 256,64,64 --reps 20 --mg
    python -m h8pcg.kbench --dims 256,64,64 --reps 50 --bilinear

Builds the clamped box operator (per-element scale 10**U(9,12), seed 5), then
times the fused direction + K*p kernel and the update kernel of one PCG
iteration (each launched `reps` times alone, captured in a HIP graph, timed
with HIP events on the launching stream) -- with --variant 7, one launch of the
resident solver running `reps` iterations -- and finally runs a plain
elementwise kernel (torch.neg, float32) over --copy-mib MiB -- it reads and
writes exactly that many bytes -- so hardware byte counters (FETCH_SIZE,
WRITE_SIZE) can be calibrated against a kernel of known traffic.  Prints one JSON line.  Run it under
``rocprofv3 --kernel-trace --stats`` or ``rocprofv3 --pmc ...``.  With --mg it
times the multigrid kernels of the FP32 iterations instead (h8_mg_time_kernels:
each kernel of the V-cycle alone, and one whole V-cycle, with its algorithmic
bytes).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json

import numpy as np


def build_operator(nx, ny, nz, seed=5, device="cuda:0"):
    from .elements import Hex8Solid3D
    from .material import D_solid3d, Material
    from .operator import GridOperator
    rng = np.random.default_rng(seed)
    s = 10 ** rng.uniform(9, 12, nx * ny * nz)
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    K0 = Hex8Solid3D().stiffness_batched(X[None], D_solid3d(Material(1.0, 0.3)), device=device)[0]
    op = GridOperator(nx, ny, nz, K0, s, device=device)
    mask = np.ones(op.n_dof, dtype=np.uint8)
    nj, ni = ny + 1, nx + 1
    k, j = np.meshgrid(np.arange(nz + 1), np.arange(nj), indexing="ij")
    clamp = (k * ni * nj + j).ravel()                  # nodes with i = 0
    mask[(3 * clamp[:, None] + np.arange(3)).ravel()] = 0
    op.set_free_mask(mask)
    return op


def time_kernels(op, precision, reps):
    from . import _lib
    ws = op.workspace(precision)
    s_e, K0 = (op.s_e, op.K0buf) if precision == 64 else (op.s_e32, op.K0_32)
    avg = (C.c_double * 2)()
    g = op.grid if precision == 64 else op.grid32
    _lib.check(_lib.lib().h8_time_kernels(C.byref(g), precision, _lib.ptr(s_e), _lib.ptr(K0),
                                          _lib.ptr(op.dinv), _lib.ptr(ws), ws.numel(), int(reps), avg,
                                          _lib.stream_handle(op.device)), "h8_time_kernels")
    return avg[0] * 1e3, avg[1] * 1e3


DEFER_X_MIN_DOF = 1000000     # the library defers x += alpha p into the FP32 sector kernel from here


def bytes_per_launch(op, precision):
    """Algorithmic bytes of the two kernels of one iteration.  Direction: read r,
    dinv, p_old, write p, q, read s_e; update: read x, p, r, q, dinv, write x, r.
    FP32 sector kernels on grids of DEFER_X_MIN_DOF+ DOFs apply the previous
    x += alpha p while they stage p (read and write x there), so the update
    streams r, q, dinv and r only."""
    w = precision // 8
    g = op.grid if precision == 64 else op.grid32
    deferred = (precision == 32 and op.n_dof >= DEFER_X_MIN_DOF and g.variant in (0, 5, 6))
    if deferred:
        return {"dir": 7 * w * op.n_dof + w * op.m, "update": 4 * w * op.n_dof}
    return {"dir": 5 * w * op.n_dof + w * op.m, "update": 7 * w * op.n_dof}


def recovery_bytes(nx, ny, nz):
    """Algorithmic HBM bytes of one recovery: nodal kernel reads u and the element
    scales once and writes 6 stresses + von Mises per node; element kernel reads
    u and the scales once and writes 6 stresses + von Mises per element."""
    n_nodes, m = (nx + 1) * (ny + 1) * (nz + 1), nx * ny * nz
    return {"nodal": 24 * n_nodes + 8 * m + 56 * n_nodes, "elements": 24 * n_nodes + 8 * m + 56 * m}


def time_recovery(nx, ny, nz, reps, device="cuda:0"):
    """Average microseconds of h8_recover_f64 (table + kernel) for the nodal and
    the element outputs, timed with events on the launching (current) stream."""
    import torch
    from .operator import recover_grid
    from .material import D_solid3d, Material
    rng = np.random.default_rng(5)
    u = torch.as_tensor(rng.normal(size=3 * (nx + 1) * (ny + 1) * (nz + 1)), device=device)
    s = 10 ** rng.uniform(9, 12, nx * ny * nz)
    D = D_solid3d(Material(1.0, 0.3))
    out = {}
    for what in ("nodal", "elements"):
        kw = dict(nodes=what == "nodal", elements=what == "elements", device=device)
        for _ in range(3):
            recover_grid(nx, ny, nz, u, (1.0, 1.0, 1.0), D, s, **kw)
        # resident scale and D: the timed calls upload nothing
        st = torch.as_tensor(s, device=device)
        Dt = torch.as_tensor(D, device=device)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            recover_grid(nx, ny, nz, u, (1.0, 1.0, 1.0), Dt, st, **kw)
        ev[1].record()
        torch.cuda.synchronize()
        out[what + "_us"] = ev[0].elapsed_time(ev[1]) * 1e3 / reps
    return out


def time_bilinear(op, reps):
    """Average microseconds of h8_elem_bilinear_f64 with u == v (the compliance
    sensitivity: one vector read) and with u != v (the adjoint product), and
    their algorithmic HBM bytes: the vectors (8 B per DOF each), w in and out
    (8 B per element each)."""
    import torch
    rng = np.random.default_rng(7)
    u = torch.as_tensor(rng.normal(size=op.n_dof), device=op.device)
    v = torch.as_tensor(rng.normal(size=op.n_dof), device=op.device)
    w = torch.as_tensor(rng.normal(size=op.m), device=op.device)
    out = {}
    for name, b in (("same", u), ("two", v)):
        for _ in range(3):
            op.element_bilinear(u, b, w)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            op.element_bilinear(u, b, w)
        ev[1].record()
        torch.cuda.synchronize()
        us = ev[0].elapsed_time(ev[1]) * 1e3 / reps
        nbytes = (1 if name == "same" else 2) * 8 * op.n_dof + 16 * op.m
        out[name] = {"us": us, "bytes": nbytes, "GBps": nbytes / us * 1e-3}
    return out


# ---------------------------------------------------------------------------
# algorithmic bytes of the multigrid-preconditioned FP32 iteration (FP32 words;
# n DOFs and m elements on level 0, a stencil level holds 243 words per node)
# ---------------------------------------------------------------------------
MG_COARSE_DOF = 128       # the library halves even axes until a level has at most this many DOFs
MG_BOTTOM_NODES = 256     # levels after the first with at most this many nodes run in one workgroup
W = 4


def mg_levels(nx, ny, nz):
    """[(ndof, nnodes)] of the library's hierarchy."""
    out = []
    while True:
        nn = (nx + 1) * (ny + 1) * (nz + 1)
        out.append((3 * nn, nn))
        r = [2 if v % 2 == 0 else 1 for v in (nx, ny, nz)]
        if 3 * nn <= MG_COARSE_DOF or r == [1, 1, 1] or len(out) == 12:
            return out
        nx, ny, nz = nx // r[0], ny // r[1], nz // r[2]


def mg_bottom_level(levels):
    lb = len(levels) - 1
    while lb > 1 and levels[lb - 1][1] <= MG_BOTTOM_NODES:
        lb -= 1
    return lb


def _restrict(nF, nC, pre0):
    return W * (nF + nC * (1 + 3 * pre0))        # r in (+ coarse dinv); coarse b (+ d, x0) out


def _prolong(nF, nC):
    return W * (3 * nF + nC)                      # coarse x, mask, x in; x out


def _st(n, N, reads, writes):
    return W * (243 * N + (reads + writes) * n)


def mg_cycle_kernels(nx, ny, nz, deg0):
    """The kernels of one V-cycle with their algorithmic bytes per launch and
    launches per cycle, keyed by the slots of h8_mg_time_kernels (deg0 =
    Chebyshev steps per level-0 smoothing pass: the first pair in slot 4, then
    deg0 - 2 steps before the residual and deg0 after the prolongation):
      3 level-0 Chebyshev step (u, s_e, b, d, dinv in; d, x out),
      4 level-0 first step pair from zero (b, dinv, s_e in; d, x out),
      5 level-0 masked residual (u, s_e, b, dinv in; r out),
      6 restriction 0 -> 1 (+ level 1's first step: r, coarse dinv in; b, d, x0 out),
      7 prolongation 1 -> 0,
      1 level-1 Chebyshev step (stencil, x, b, d, dinv in; d, x out),
      9 level-1 masked residual (stencil, x, b, dinv in; r out),
      8 the single-workgroup kernel of the small levels."""
    levels = mg_levels(nx, ny, nz)
    m = nx * ny * nz
    n0 = levels[0][0]
    lb = mg_bottom_level(levels)
    out = {
        3: {"kernel": "k_sector<float,16,16,MODE_CHEB> (level-0 Chebyshev step)", "bytes": W * (6 * n0 + m),
            "calls": 2 * int(deg0) - 2},
        4: {"kernel": "k_sector<float,16,16,MODE_CHEB0> (level-0 first step pair)", "bytes": W * (4 * n0 + m),
            "calls": 1},
        5: {"kernel": "k_sector<float,16,16,MODE_MGRES> (level-0 masked residual)", "bytes": W * (4 * n0 + m),
            "calls": 1},
    }
    if len(levels) > 1:
        n1, N1 = levels[1]
        out[6] = {"kernel": "k_mg_restrict (0 -> 1)", "bytes": _restrict(n0, n1, 1 if lb > 1 else 0), "calls": 1}
        out[7] = {"kernel": "k_mg_prolong (1 -> 0)", "bytes": _prolong(n0, n1), "calls": 1}
        if lb > 1:
            out[1] = {"kernel": "k_mg_cheb_st (level-1 Chebyshev step)", "bytes": _st(n1, N1, 4, 2), "calls": 3}
            out[9] = {"kernel": "k_mg_resid_st (level-1 masked residual)", "bytes": _st(n1, N1, 3, 1), "calls": 1}
        out[8] = {"kernel": "k_mg_bottom (levels %d..%d in one workgroup)" % (lb, len(levels) - 1),
                  "bytes": _bottom_bytes(levels, lb), "calls": 1}
    return out


def _bottom_bytes(levels, lb):
    b = 0
    for li in range(lb, len(levels)):
        n, N = levels[li]
        if li == len(levels) - 1:
            b += W * (n * n + 2 * n)
            break
        nc = levels[li + 1][0]
        b += W * 4 * n + _st(n, N, 4, 2) + _st(n, N, 3, 1) + _restrict(n, nc, 0)
        b += _prolong(n, nc) + _st(n, N, 3, 2) + _st(n, N, 4, 2)
    return b


def mg_iteration_bytes(nx, ny, nz, deg0):
    """Algorithmic HBM bytes of one FP32 iteration of the multigrid-preconditioned
    refined solve, the kernels as the library runs them:
      PCG: direction + product (z, p_old, s_e in; p, q out), update (x, r, p,
           q, dinv in; x, r out);
      level 0: first step pair, deg0 - 2 Chebyshev steps, masked residual,
           restriction, prolongation, deg0 Chebyshev steps (mg_cycle_kernels);
      stencil level l (1 <= l < the bottom level): first step in the
           restriction into it, one Chebyshev step and the masked residual,
           restriction, prolongation, two Chebyshev steps;
      the small levels as the single-workgroup kernel reads them.
    Returns (bytes per iteration, bytes per FP64 residual replacement, levels);
    a replacement folds x into u (u, x in; u, x out) and forms r = f - K u in
    FP64 (u, f, dinv, s_e in; r32 out)."""
    levels = mg_levels(nx, ny, nz)
    m = nx * ny * nz
    n0 = levels[0][0]
    lb = mg_bottom_level(levels)
    pcg = W * (4 * n0 + m) + W * 7 * n0
    cyc = 0
    if len(levels) == 1:
        cyc = W * (n0 * n0 + 2 * n0)
    else:
        cyc = sum(k["bytes"] * k["calls"] for key, k in mg_cycle_kernels(nx, ny, nz, deg0).items() if key not in (1, 9))
        for li in range(1, lb):
            n, N = levels[li]
            nc = levels[li + 1][0]
            cyc += _st(n, N, 4, 2) + _st(n, N, 3, 1) + _st(n, N, 3, 2) + _st(n, N, 4, 2)
            if li + 1 < len(levels):
                cyc += _restrict(n, nc, 1 if li + 1 < lb else 0) + _prolong(n, nc)
    rep = 8 * 2 * n0 + 4 * 2 * n0 + 8 * 3 * n0 + 8 * m + 4 * n0
    return pcg + cyc, rep, levels


def time_mg(op, reps):
    """The multigrid kernels (h8_mg_time_kernels, HIP events on the launching
    stream) with their algorithmic bytes: {slot: {kernel, bytes, calls,
    avg_us, GBps}}, plus the whole V-cycle and the level hierarchy."""
    from . import _lib
    ws, mws = op.workspace(32), op.mg_workspace()
    avg = (C.c_double * _lib.MG_TIMES)()
    info = (C.c_int64 * 4)()
    _lib.check(_lib.lib().h8_mg_time_kernels(C.byref(op.grid32), _lib.ptr(op.s_e), _lib.ptr(op.K0buf),
                                             _lib.ptr(op.s_e32), _lib.ptr(op.K0_32), _lib.ptr(op.dinv),
                                             _lib.ptr(ws), ws.numel(), _lib.ptr(mws), mws.numel(), int(reps), avg,
                                             info, _lib.stream_handle(op.device)), "h8_mg_time_kernels")
    kern = mg_cycle_kernels(op.nx, op.ny, op.nz, int(info[3]))
    kern[0] = {"kernel": "k_sector<float,16,16,MODE_PLAIN> (level-0 product y = K x)",
               "bytes": W * (2 * op.n_dof + op.m), "calls": 0}
    for slot, k in kern.items():
        k["avg_us"] = 1e3 * avg[slot]
        k["GBps"] = k["bytes"] / k["avg_us"] * 1e-3 if k["avg_us"] > 0 else None
    return {"levels": int(info[0]), "level1_dof": int(info[1]), "level1_nodes": int(info[2]),
            "mg_degree": int(info[3]), "vcycle_us": 1e3 * avg[2], "kernels": kern}


def main(argv=None):
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--dims", default="256,64,64")
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--precision", default="32,64")
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--layers", type=int, default=0, help="node layers per block of the sector kernel (0 = auto)")
    ap.add_argument("--arith", type=int, default=0, help="FP32 element-product arithmetic (0 = library, 32, 64)")
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--recovery", action="store_true", help="time the stress recovery kernels instead")
    ap.add_argument("--mg", action="store_true", help="time the multigrid kernels instead")
    ap.add_argument("--bilinear", action="store_true", help="time the element bilinear form instead")
    a = ap.parse_args(argv)
    nx, ny, nz = (int(v) for v in a.dims.split(","))
    if a.recovery:
        r = time_recovery(nx, ny, nz, a.reps)
        b = recovery_bytes(nx, ny, nz)
        r.update({"dims": [nx, ny, nz], "reps": a.reps, "nodal_bytes": b["nodal"], "elements_bytes": b["elements"],
                  "nodal_GBps": b["nodal"] / r["nodal_us"] * 1e-3,
                  "elements_GBps": b["elements"] / r["elements_us"] * 1e-3})
        print(json.dumps(r), flush=True)
        return
    op = build_operator(nx, ny, nz)
    op.grid.variant = a.variant
    op.grid.layers = a.layers
    op.grid32.variant = a.variant
    op.grid32.layers = a.layers
    op.set_arith(a.arith)
    out = {"dims": [nx, ny, nz], "n_dof": op.n_dof, "elements": op.m, "reps": a.reps}
    if a.bilinear:
        out["bilinear"] = time_bilinear(op, a.reps)
    if a.mg:
        out["mg"] = time_mg(op, a.reps)
    for prec in ((int(p) for p in a.precision.split(",")) if not (a.mg or a.bilinear) else ()):
        d_us, u_us = time_kernels(op, prec, a.reps)
        b = bytes_per_launch(op, prec)
        if a.variant == 7:
            # the resident solver: one launch of `reps` iterations; d_us is per iteration
            it_b = b["dir"] + b["update"]
            out[f"fp{prec}"] = {"iteration_us": d_us, "launch_us": d_us * a.reps, "iterations_per_launch": a.reps,
                                "iteration_bytes": it_b, "iteration_GBps": it_b / d_us * 1e-3}
            continue
        out[f"fp{prec}"] = {"dir_us": d_us, "update_us": u_us, "dir_bytes": b["dir"], "update_bytes": b["update"],
                            "dir_GBps": b["dir"] / d_us * 1e-3, "update_GBps": b["update"] / u_us * 1e-3}
    if a.copy_mib > 0:
        n = a.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=op.device)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        out["calibration_kernel"] = "neg (float32)"
        out["calibration_bytes_each_way"] = 4 * n
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
