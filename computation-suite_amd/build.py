"""Build libh8pcg.so (HIP, gfx950) in-tree: computation-suite_amd/h8pcg/libh8pcg.so.

    python computation-suite_amd/build.py [--arch gfx950] [--verbose]

All device code is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); hipcc's host pass warns
that it does not know that feature, which is harmless.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "h8pcg", "libh8pcg.so")
SOURCES = ["h8_lib.hip"]
HEADERS = ["h8_device.h", "h8_node.h", "h8_stream.h", "h8_mfma.h", "h8_batched.h", "h8_sector.h", "h8_resident.h", "h8_recovery.h", "h8_topopt.h", "h8_constrained.h", "h8_mg.h", "h8_hex20.h"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(arch: str = "gfx950", verbose: bool = False, extra: list[str] | None = None) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [hipcc(), f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
           "-I", os.path.join(HERE, "..", "include"), "-I", CSRC,
           "-o", OUT + ".tmp", *srcs]
    if extra:
        cmd[1:1] = extra
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    if verbose and res.stderr.strip():
        # keep real diagnostics visible, drop the known host-pass feature notice
        lines = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    os.replace(OUT + ".tmp", OUT)
    return OUT


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.arch, a.verbose))
    return 0


if __name__ == "__main__":
    sys.exit(main())
