"""Pins the oracle's constraint elimination to the reference's own outputs:
the reduced map of an overlapping constraint set, and the constrained solve of
a tied, multi-point-constrained cantilever with a prescribed displacement."""
import os

import numpy as np
import pytest

from oracle import h8_oracle as o

GOLD = os.path.join(os.path.dirname(__file__), "golden", "constraint_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_reduce_matches_reference(gold):
    rows = [{3: 1.0, 7: -1.0}, {7: 2.0, 11: -1.0, 5: 0.5}, {11: 1.0, 2: 1.0}, {20: 1.0, 21: -1.0}]
    slave, mat, g = o.reduce_constraints(rows, [0.0, 1.5, -0.25, 0.0], {2: 0.1})
    assert list(slave) == list(gold["reduce/slave"])
    masters = list(gold["reduce/masters"])
    for i, s in enumerate(slave):
        got = np.zeros(len(masters))
        for m, c in mat[int(s)].items():
            got[masters.index(m)] = c
        assert np.allclose(got, gold["reduce/mat"][i], rtol=0, atol=1e-15)
        assert np.isclose(g[int(s)], gold["reduce/g"][i], rtol=0, atol=1e-15)


def constrained_case(gold):
    nx, ny, nz = (int(v) for v in gold["solve/dims"])
    c = o.cantilever(nx, ny, nz, E_elem=gold["solve/E"])
    fixed = list(c["fixed_dofs"]) + [3 * int(gold["solve/mid"])]
    vals = [0.0] * len(c["fixed_dofs"]) + [1e-4]
    cons = [(np.array([3 * int(a) + 2, 3 * int(b) + 2]), np.array([1.0, -1.0]), 0.0)
            for a, b in zip(gold["solve/y0"], _partners(c["nodes"], gold["solve/y0"], ny))]
    a, b, cc = (int(v) for v in gold["solve/abc"])
    cons.append((np.array([3 * a, 3 * b, 3 * cc + 1]), np.array([1.0, -2.0, 0.5]), 2e-5))
    return c, fixed, vals, cons


def _partners(nodes, y0, ny):
    out = []
    for n in y0:
        x = nodes[int(n)] + np.array([0.0, ny, 0.0])
        out.append(int(np.argmin(np.linalg.norm(nodes - x, axis=1))))
    return out


def test_constrained_solve_matches_reference(gold):
    c, fixed, vals, cons = constrained_case(gold)
    assert np.allclose(c["F"], gold["solve/F"], rtol=0, atol=1e-6)
    U = o.solve_constrained(c["K"], c["F"], fixed, vals, cons)
    ref = gold["solve/U"]
    assert np.linalg.norm(U - ref) / np.linalg.norm(ref) < 1e-10
    for d, co, v in cons:
        assert abs(co @ U[d] - v) < 1e-12 * max(1.0, np.abs(U).max())
