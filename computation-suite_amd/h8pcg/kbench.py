"""Kernel micro-benchmark for profiling the two PCG iteration kernels.

    python -m h8pcg.kbench --dims 256,64,64 --reps 50 [--precision 32,64] [--variant V] [--layers L] [--copy-mib 256]

Builds the clamped box operator (per-element scale 10**U(9,12), seed 5), then
times the fused direction + K*p kernel and the update kernel of one PCG
iteration (each launched `reps` times alone, captured in a HIP graph, timed
with HIP events on the launching stream) -- with --variant 7, one launch of the
resident solver running `reps` iterations -- and finally runs a plain
elementwise kernel (torch.neg, float32) over --copy-mib MiB -- it reads and
writes exactly that many bytes -- so hardware byte counters (FETCH_SIZE,
WRITE_SIZE) can be calibrated against a kernel of known traffic.  Prints one JSON line.  Run it under
``rocprofv3 --kernel-trace --stats`` or ``rocprofv3 --pmc ...``.
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


def bytes_per_launch(op, precision):
    w = precision // 8
    return {"dir": 5 * w * op.n_dof + w * op.m, "update": 7 * w * op.n_dof}


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
    a = ap.parse_args(argv)
    nx, ny, nz = (int(v) for v in a.dims.split(","))
    op = build_operator(nx, ny, nz)
    op.grid.variant = a.variant
    op.grid.layers = a.layers
    op.grid32.variant = a.variant
    op.grid32.layers = a.layers
    op.set_arith(a.arith)
    out = {"dims": [nx, ny, nz], "n_dof": op.n_dof, "elements": op.m, "reps": a.reps}
    for prec in (int(p) for p in a.precision.split(",")):
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
