"""CPU tests of the host side: the C-ABI library loads and exports every symbol
the header declares, and the FESystem mirror's host logic (meshes, node
selection, loads, BC bookkeeping, validation errors) matches the oracle."""
import os
import re

import numpy as np
import pytest

from oracle import h8_oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "h8pcg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(h8_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from h8pcg import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the binding knows a signature for each of them
    assert not [n for n in names if n not in _lib.SIGNATURES]


def test_library_non_device_calls():
    from h8pcg import _lib
    L = _lib.lib()
    assert L.h8_version() >= 1
    g = _lib.H8Grid(128, 32, 32, 0)
    n = 3 * 129 * 33 * 33
    b64 = L.h8_pcg_workspace_bytes(g, 64)
    b32 = L.h8_pcg_workspace_bytes(g, 32)
    assert b64 >= 4 * 8 * n and b32 >= 6 * 4 * n
    bad = _lib.H8Grid(0, 1, 1, 0)
    assert L.h8_pcg_workspace_bytes(bad, 64) == 0
    # argument validation happens before any device work
    rc = L.h8_spmv_f64(bad, None, None, None, None, None)
    assert rc == _lib.H8_ERR_ARG and "nx, ny, nz" in _lib.last_error()
    with pytest.raises(ValueError):
        _lib.check(rc)
    # the multigrid smoothing degree: 0 (library choice) or 2..6
    for deg, ok in ((0, True), (2, True), (6, True), (1, False), (7, False), (-1, False)):
        gd = _lib.H8Grid(128, 32, 32, 0, 0, 0, deg)
        assert (L.h8_mg_workspace_bytes(gd) > 0) == ok, deg
        if not ok:
            rc = L.h8_spmv_f64(gd, None, None, None, None, None)
            assert rc == _lib.H8_ERR_ARG and "mg_degree" in _lib.last_error()


def test_box_mesh_matches_oracle_and_structure():
    import h8pcg
    m = h8pcg.box_mesh(7.0, 3.0, 2.0, 7, 3, 4)
    nodes, elems = o.box_mesh(7.0, 3.0, 2.0, 7, 3, 4)
    assert np.array_equal(m.elements, elems)
    assert np.abs(m.nodes - nodes).max() < 1e-14
    nx, ny, nz, X0 = h8pcg.structured_grid(m)
    assert (nx, ny, nz) == (7, 3, 4)
    # a plain Mesh with the same arrays (e.g. built by the reference) is recognised too
    plain = h8pcg.Mesh(nodes.copy(), elems.copy(), 3)
    assert h8pcg.structured_grid(plain)[:3] == (7, 3, 4)
    # a distorted interior node breaks the shared-K0 structure
    plain.nodes[len(nodes) // 2] += 0.1
    assert h8pcg.structured_grid(plain) is None


def test_loads_and_dofs_match_oracle():
    import h8pcg
    mesh = h8pcg.box_mesh(20.0, 4.0, 4.0, 20, 4, 4)
    fs = h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D())
    fs.add_consistent_facet_load(h8pcg.Hex8Solid3D(), h8pcg.boundary_faces(mesh, 0, 20.0), 2, -1e6 / 16)
    c = o.cantilever(20, 4, 4)
    assert np.abs(fs.F - c["F"]).max() < 1e-9
    fs.fix_dofs(mesh.nodes_on_plane(0, 0.0), ["ux", "uy", "uz"])
    assert sorted(fs.fixed_dofs) == sorted(c["fixed_dofs"])
    free, fixed = o.partition(fs.n_dof, c["fixed_dofs"])
    assert np.array_equal(fs.free_dofs, free) and np.array_equal(fs.fixed_dofs_array, fixed)
    fs2 = h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D())
    fs2.add_nodal_force(mesh.nodes_on_plane(0, 20.0), "uz", -1e6)
    c2 = o.cantilever(20, 4, 4, load="nodal")
    assert np.abs(fs2.F - c2["F"]).max() < 1e-9


def test_validation_errors():
    import h8pcg
    mesh = h8pcg.box_mesh(2.0, 1.0, 1.0, 2, 1, 1)
    with pytest.raises(ValueError, match="backend"):
        h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D(), backend="scipy")
    fs = h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D())
    with pytest.raises(ValueError, match="no nodes selected"):
        fs.fix_dofs([], [0])
    with pytest.raises(ValueError, match=r"Hex8Solid3D has 3 DOF\(s\) per node.*0\.\.2"):
        fs.fix_dofs([0], [3])
    with pytest.raises(ValueError, match="outside the mesh"):
        fs.add_nodal_force([99], 0, 1.0)
    with pytest.raises(ValueError, match="unknown node set"):
        fs.fix_dofs("tip", [0])
    mesh.select_nodes(x=2.0, name="tip")
    fs.fix_dofs("tip", ["ux"])
    assert len(fs.fixed_dofs) == 4
    with pytest.raises(RuntimeError, match="assemble_stiffness"):
        fs.solve_static()
    with pytest.raises(ValueError, match="unknown method"):
        fs.assemble_stiffness(np.eye(6), method="bogus")
    with pytest.raises(RuntimeError, match="all zeros"):
        fs.solve_static()
    # fe/tests/test_vectorized_assembly.py:98-103
    with pytest.raises(ValueError, match="vectorized=True only supports"):
        fs.assemble_stiffness(np.eye(6), method="reduced", vectorized=True)


def test_coefficient_factoring():
    from h8pcg import coefficients, D_solid3d, Material
    E = np.array([1.0, 3.0, 7.0])
    D_ref, s = coefficients.factor_scales(coefficients.by_element([D_solid3d(Material(e, 0.3)) for e in E]))
    assert np.allclose(s * D_ref[0, 0], [D_solid3d(Material(e, 0.3))[0, 0] for e in E])
    with pytest.raises(ValueError, match="scalar multiples"):
        coefficients.factor_scales(coefficients.by_element([D_solid3d(Material(1, 0.3)), D_solid3d(Material(1, 0.2))]))


def test_constraint_reduction_matches_reference_fixture():
    """The product's grouped constraint reduction (h8pcg.constraints) gives the
    reference's reduced map (tests/golden/constraint_golden.npz)."""
    import numpy as np
    from h8pcg.constraints import reduce_constraints
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "constraint_golden.npz"))
    rows = [{3: 1.0, 7: -1.0}, {7: 2.0, 11: -1.0, 5: 0.5}, {11: 1.0, 2: 1.0}, {20: 1.0, 21: -1.0}]
    slave, mat, gg = reduce_constraints(rows, [0.0, 1.5, -0.25, 0.0], {2: 0.1})
    ref = {int(s): i for i, s in enumerate(g["reduce/slave"])}
    assert sorted(slave.tolist()) == sorted(ref)
    masters = list(g["reduce/masters"])
    for s in slave:
        got = np.zeros(len(masters))
        for m, c in mat[int(s)].items():
            got[masters.index(m)] = c
        assert np.allclose(got, g["reduce/mat"][ref[int(s)]], rtol=0, atol=1e-15)
        assert np.isclose(gg[int(s)], g["reduce/g"][ref[int(s)]], rtol=0, atol=1e-15)
    import pytest
    with pytest.raises(ValueError, match="inconsistent"):
        reduce_constraints([{1: 1.0, 2: -1.0}, {1: 2.0, 2: -2.0}], [0.0, 1.0], {})


def test_add_constraint_argument_checks():
    import numpy as np
    import pytest
    import h8pcg
    mesh = h8pcg.box_mesh(2.0, 1.0, 1.0, 2, 1, 1)
    fs = h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D(), backend="hip", device="cuda")
    with pytest.raises(ValueError, match="no terms"):
        fs.add_constraint([])
    with pytest.raises(ValueError, match="twice"):
        fs.add_constraint([(1, "ux", 1.0), (1, 0, -1.0)])
    with pytest.raises(ValueError, match="all coefficients are zero"):
        fs.add_constraint([(1, "ux", 0.0), (2, "ux", 0.0)])
    fs.add_constraint([(1, "ux", 1.0), (2, "ux", -1.0)], value=0.5)
    assert len(fs._constraints) == 1 and fs._constraints[0][2] == 0.5
    fs.clear_constraints()
    assert fs._constraints == ()
    with pytest.raises(ValueError, match="no master node"):
        fs.tie([0], [5], [0], offset=[10.0, 0.0, 0.0])


def test_rigid_null_basis_spans_null_space_of_Kff():
    """The rigid motions left free by the supports span the null space of Kff
    (checked against the oracle's assembled K and the dimension of its null
    space by eigenvalues, the test fe/solver.py:1583-1678 applies)."""
    import h8pcg
    mesh = h8pcg.box_mesh(3.0, 1.0, 1.0, 3, 2, 2)
    nodes, elems = o.box_mesh(3.0, 1.0, 1.0, 3, 2, 2)
    K = o.assemble_csr(nodes, elems, o.d_solid3d(1.0, 0.3)).toarray()

    def null_dim(fixed):
        free, _ = o.partition(K.shape[0], np.asarray(fixed, dtype=np.int64))
        lam = np.linalg.eigvalsh(K[np.ix_(free, free)])
        return int(np.sum(np.abs(lam) < 1e-8 * np.abs(lam).max())), free

    x0 = mesh.nodes_on_plane(0, 0.0)
    cases = [([], []), (x0, [0]), (x0, [0, 1, 2]), ([0], [0, 1, 2]), (mesh.select_nodes(x=0.0, y=0.0), [0, 1, 2])]
    dims = []
    for nodes_fixed, comps in cases:
        fs = h8pcg.FESystem(mesh, h8pcg.Hex8Solid3D())
        if len(nodes_fixed):
            fs.fix_dofs(nodes_fixed, comps)
        Q = fs._rigid_null_basis()
        r = 0 if Q is None else Q.shape[1]
        nd, free = null_dim(fs.fixed_dofs_array)
        assert r == nd
        dims.append(r)
        if Q is not None:
            assert np.allclose(Q.T @ Q, np.eye(r), atol=1e-12)
            assert np.all(Q[fs.fixed_dofs_array] == 0.0)
            assert np.max(np.abs(K[np.ix_(free, free)] @ Q[free])) < 1e-12 * np.abs(K).max()
        assert fs._rigid_null_basis() is Q          # cached per fixed set
    assert dims == [6, 3, 0, 3, 1]


def test_hex8_face_tables():
    """fe/tests/test_facet_loads.py:101-119: on every face's natural-coordinate
    map the element's shape functions vanish at the other four nodes and equal
    the bilinear facet's at the face's nodes, in the table's order."""
    import h8pcg
    from h8pcg import facet_loads
    elem = h8pcg.Hex8Solid3D()
    rng = np.random.default_rng(8)
    facets = facet_loads.list_facets(elem)
    assert len(facets) == 6
    for local_idx, family, natcoord_map in facets:
        others = [i for i in range(8) if i not in local_idx]
        shape_fn = facet_loads._FACET_SHAPE_FNS[family]
        for _ in range(8):
            a, b = rng.uniform(-0.9, 0.9, 2)
            N, _ = elem.shape_and_derivs(natcoord_map(a, b))
            Nf, _ = shape_fn((a, b))
            assert np.max(np.abs(N[others])) < 1e-10
            assert np.max(np.abs(N[list(local_idx)] - Nf)) < 1e-9
    # the outward normal of each face (cross product of the facet's tangents
    # in table order) points away from the element centre
    cube = np.array(facet_loads._HEX8_NAT)
    for local_idx, _, nat in facets:
        P = np.array([nat(a, b) for a, b in [(-1, -1), (1, -1), (-1, 1)]])
        n = np.cross(P[1] - P[0], P[2] - P[0])
        assert np.dot(n, cube[list(local_idx)].mean(axis=0)) > 0
    with pytest.raises(ValueError, match="no facet table"):
        facet_loads.list_facets(h8pcg.Hex8SolidBbar())
