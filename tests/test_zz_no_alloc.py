"""The FP64 Jacobi-PCG entry point that allocates nothing (h8_pcg_f64_noalloc,
workspace in the library's static device scratch) against h8_pcg_f64 with a
caller workspace: bitwise equal solutions and iteration counts on two calls in
a row, the second on a smaller grid, for the one-launch resident solver and
for the streaming kernels (direct launches, no captured graph); grids beyond
the scratch are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(h8, nx, ny, nz, seed):
    mesh = h8.box_mesh(float(nx), float(ny), float(nz), nx, ny, nz)
    fs = h8.FESystem(mesh, h8.Hex8Solid3D(), backend="hip", device="cuda:0")
    E = 10 ** np.random.default_rng(seed).uniform(9, 11, nx * ny * nz)
    fs.assemble_stiffness(h8.D_solid3d(h8.Material(E=1.0, nu=0.3)), element_scale=E)
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), [0, 1, 2])
    fs.add_consistent_facet_load(h8.Hex8Solid3D(), h8.boundary_faces(mesh, 0, float(nx)), 2, -1e6 / (ny * nz))
    return fs


def _noalloc(op, b, tol):
    import torch
    from h8pcg import _lib
    x = torch.empty_like(b)
    st = _lib.H8PcgStats()
    with torch.cuda.device(op.device):
        rc = _lib.lib().h8_pcg_f64_noalloc(C.byref(op.grid), _lib.ptr(op.s_e), _lib.ptr(op.K0buf), _lib.ptr(op.dinv),
                                           _lib.ptr(b), _lib.ptr(x), 0, float(tol), 10 * op.n_dof, C.byref(st),
                                           _lib.stream_handle(op.device))
    _lib.check(rc, "h8_pcg_f64_noalloc")
    return x, st.as_dict()


@pytest.mark.parametrize("variant", [0, 1])
def test_noalloc_pcg_matches_workspace_pcg(gpu, variant):
    import h8pcg
    for dims, seed in (((48, 12, 12), 1), ((16, 4, 4), 2)):
        fs = _system(h8pcg, *dims, seed)
        op = fs.operator
        op.set_variant(variant)
        op.set_free_mask(fs._free_mask())
        b, _ = fs._dirichlet_rhs(op, fs.F)
        x_ws, st_ws = op.pcg(b, tol=1e-10)
        x_na, st_na = _noalloc(op, b, 1e-10)
        print("[gpu-test] noalloc", variant, dims, st_ws["iterations"], st_na["iterations"], st_na["launches"],
              flush=True)
        assert st_na["iterations"] == st_ws["iterations"] and st_na["status"] == 0
        assert st_na["elapsed_ms"] == -1.0
        assert np.array_equal(x_na.cpu().numpy(), x_ws.cpu().numpy())


def test_noalloc_refuses_grids_beyond_the_scratch(gpu):
    import torch
    from h8pcg import _lib
    g = _lib.H8Grid(256, 64, 64, 0, 0, 0)
    assert _lib.lib().h8_pcg_workspace_bytes(C.byref(g), 64) > _lib.SCRATCH_BYTES
    t = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    st = _lib.H8PcgStats()
    rc = _lib.lib().h8_pcg_f64_noalloc(C.byref(g), _lib.ptr(t), _lib.ptr(t), _lib.ptr(t), _lib.ptr(t), _lib.ptr(t),
                                       0, 1e-8, 10, C.byref(st), _lib.stream_handle("cuda:0"))
    assert rc == _lib.H8_ERR_WORKSPACE
    assert "static scratch" in _lib.last_error()
    g = _lib.H8Grid(128, 32, 32, 0, 0, 0)
    assert _lib.lib().h8_pcg_workspace_bytes(C.byref(g), 64) <= _lib.SCRATCH_BYTES
