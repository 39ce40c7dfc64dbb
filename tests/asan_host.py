"""AddressSanitizer + UndefinedBehaviorSanitizer build of the library's HOST code
and the driver that exercises it without a GPU (test infrastructure).

The build compiles csrc/h8_lib.hip with the sanitizers on the host pass only
(each -fsanitize= follows -Xarch_host; the gfx950 device code is built as
usual) into _build/asan/libh8pcg_asan.so.  The driver runs in a child process
with the ASan runtime preloaded (Python itself is not instrumented) and calls
every entry point whose work before the first HIP call is host code: the
workspace and plan sizing, the grid and instance-table validation, and the
argument checks and error-message formatting of the solver entry points with
invalid arguments.  Any ASan or UBSan report fails the run.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "computation-suite_amd", "csrc")
OUT_DIR = os.path.join(ROOT, "_build", "asan")
OUT = os.path.join(OUT_DIR, "libh8pcg_asan.so")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    return None


def asan_runtime():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def sources():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [os.path.join(ROOT, "include", "h8pcg.h")]


def build():
    """Builds OUT unless it is newer than every source; returns its path."""
    if os.path.exists(OUT) and os.path.getmtime(OUT) > max(os.path.getmtime(s) for s in sources()):
        return OUT
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = [hipcc(), "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-o", OUT + ".tmp", os.path.join(CSRC, "h8_lib.hip")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("sanitizer build failed:\n" + res.stdout + res.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


# The child process: ctypes calls into the sanitizer build.  Returns through
# its exit status; prints one line per check.
DRIVER = r'''
import ctypes as C, sys
L = C.CDLL(sys.argv[1])

class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("variant", C.c_int32),
                ("layers", C.c_int32), ("arith", C.c_int32), ("mg_degree", C.c_int32)]

class Inst(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nrhs", C.c_int32),
                ("node_off", C.c_int64), ("elem_off", C.c_int64), ("vec_off", C.c_int64)]

L.h8_last_error.restype = C.c_char_p
for name in ("h8_pcg_workspace_bytes", "h8_mg_workspace_bytes", "h8_batched_workspace_bytes",
             "h8_oc_workspace_bytes", "h20_pcg_workspace_bytes"):
    getattr(L, name).restype = C.c_size_t
L.h8_pcg_workspace_bytes.argtypes = [C.POINTER(Grid), C.c_int]
L.h8_mg_workspace_bytes.argtypes = [C.POINTER(Grid)]
L.h8_batched_workspace_bytes.argtypes = [C.POINTER(Inst), C.c_int32, C.c_int64]
L.h8_oc_workspace_bytes.argtypes = [C.c_int64]
L.h20_pcg_workspace_bytes.argtypes = [C.c_int32] * 3

def err():
    return L.h8_last_error().decode()

checks = 0
def ok(cond, what):
    global checks
    checks += 1
    if not cond:
        print("FAIL", what, "|", err(), flush=True)
        sys.exit(3)

ok(L.h8_version() > 0, "version")
# workspace and plan sizing over grid shapes, variants (not the resident
# planner, which asks the device for its CU count) and precisions
prev = 0
for dims in [(1, 1, 1), (2, 2, 2), (5, 3, 2), (20, 4, 4), (33, 17, 9), (128, 32, 32), (256, 64, 64), (300, 200, 150)]:
    for v in range(0, 7):
        for prec in (32, 64):
            b = L.h8_pcg_workspace_bytes(C.byref(Grid(*dims, v, 0, 0)), prec)
            ok(b > 0, f"pcg workspace {dims} v{v} p{prec}")
    mg = L.h8_mg_workspace_bytes(C.byref(Grid(*dims, 0, 0, 0)))
    print("mg workspace", dims, mg, flush=True)
    ok(L.h20_pcg_workspace_bytes(*dims) > 0, f"hex20 workspace {dims}")
ok(L.h8_mg_workspace_bytes(C.byref(Grid(7, 3, 3, 0, 0, 0))) == 0, "mg on a grid that does not coarsen")
for deg, want in ((0, True), (2, True), (6, True), (1, False), (7, False), (-3, False)):
    ok((L.h8_mg_workspace_bytes(C.byref(Grid(16, 4, 4, 0, 0, 0, deg))) > 0) == want, f"mg_degree {deg}")
# invalid grids: zero size, negative layers, unknown variant/arith, too many DOFs, NULL
for bad, frag in [((0, 4, 4, 0, 0, 0), "nx, ny, nz"), ((4, 4, 4, 99, 0, 0), "variant"),
                  ((4, 4, 4, 0, -1, 0), "layers"), ((4, 4, 4, 0, 0, 16), "arith"),
                  ((2000, 2000, 2000, 0, 0, 0), "too large"), ((-5, 1, 1, 0, 0, 0), "nx, ny, nz")]:
    ok(L.h8_pcg_workspace_bytes(C.byref(Grid(*bad)), 64) == 0, f"bad grid {bad}")
    ok(frag in err(), f"message for {bad}")
ok(L.h8_pcg_workspace_bytes(None, 64) == 0 and "NULL" in err(), "NULL grid")
ok(L.h8_pcg_workspace_bytes(C.byref(Grid(4, 4, 4, 0, 0, 0)), 16) == 0, "bad precision")
# batched instance tables: valid (uneven, many cases), then invalid entries
insts = []
node = elem = vec = 0
for k in range(40):
    nx = 8 + 4 * (k % 11)
    ny = nz = max(1, nx // 4)
    nrhs = 1 + k % 16
    insts.append(Inst(nx, ny, nz, nrhs, node, elem, vec))
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    node += nn
    elem += nx * ny * nz
    vec += 3 * nn * nrhs
arr = (Inst * len(insts))(*insts)
b = L.h8_batched_workspace_bytes(arr, len(insts), vec)
ok(b > 32 * vec, "batched workspace")
ok(L.h8_batched_workspace_bytes(arr, 0, vec) == 0, "no instances")
ok(L.h8_batched_workspace_bytes(None, 3, vec) == 0, "NULL instances")
for field, val in [("nx", 0), ("nrhs", 0), ("ny", -3)]:
    bad = (Inst * 3)(*insts[:3])
    setattr(bad[1], field, val)
    ok(L.h8_batched_workspace_bytes(bad, 3, vec) == 0 and "instance 1" in err(), f"bad instance {field}")
ok(L.h8_batched_workspace_bytes(arr, len(insts), -1) == 0 and "total_vec" in err(), "negative total_vec")
huge = (Inst * 1)(Inst(200, 200, 200, 500000, 0, 0, 0))
ok(L.h8_batched_workspace_bytes(huge, 1, 1 << 62) == 0 and "too large" in err(), "batch beyond int tables")
ok(L.h8_oc_workspace_bytes(1000) > 0, "oc workspace")
# solver entry points with NULL buffers: the argument checks (host) answer
# before any device work
L.h8_refine_solve_mg.argtypes = [C.POINTER(Grid)] + [C.c_void_p] * 7 + [C.c_double, C.c_int64, C.c_double, C.c_int,
                                 C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
rc = L.h8_refine_solve_mg(C.byref(Grid(8, 4, 4, 0, 0, 0)), *([None] * 7), 1e-8, 10, 0.1, 0, None, 0, None, 0,
                          None, None)
ok(rc == 1 and "NULL" in err(), "refine_solve_mg NULL")
L.h8_pcg_f64.restype = C.c_int
rc = L.h8_pcg_f64(C.byref(Grid(0, 4, 4, 0, 0, 0)), *([None] * 6), C.c_double(1e-8), C.c_int64(10),
                  None, C.c_size_t(0), None, None)
ok(rc == 1, "pcg_f64 bad grid")
print("host checks passed:", checks, flush=True)
'''


def run(lib):
    """Runs DRIVER on `lib` under the sanitizer runtime; returns (rc, output)."""
    env = dict(os.environ)
    rt = asan_runtime()
    env["LD_PRELOAD"] = rt + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:exitcode=86"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=87"
    env["HIP_VISIBLE_DEVICES"] = env.get("HIP_VISIBLE_DEVICES", "")
    res = subprocess.run([sys.executable, "-c", DRIVER, lib], capture_output=True, text=True, env=env, timeout=600)
    return res.returncode, res.stdout + res.stderr
