"""Per-element coefficients (the by_element subset of fe/coefficients.py:47-171).

``by_element(values)`` gives one D (6x6) per element.  The device operator
stores one reference element matrix K0 and a per-element scale s_e, so the
D's must be scalar multiples of one matrix (the usual case: per-element
Young's modulus with a common Poisson ratio)."""
from __future__ import annotations

import numpy as np


class Coefficient:
    def __init__(self, values):
        self.values = values

    def __repr__(self):
        return "Coefficient(by_element)"


def by_element(values):
    """Coefficient with one value per element, in connectivity order."""
    return Coefficient(list(values) if not isinstance(values, np.ndarray) else values)


def has_coefficient(*objs):
    return any(isinstance(o, Coefficient) for o in objs)


def factor_scales(D):
    """(D_ref (6,6), s (m,)) with D_e = s_e * D_ref for a by_element coefficient.

    Raises ValueError when the element matrices are not proportional to one
    another (then no single K0 exists)."""
    vals = D.values
    arr = np.asarray(vals, dtype=float)
    if arr.ndim == 1:                    # one scalar multiplier per element
        raise ValueError("by_element values must be 6x6 matrices")
    if arr.ndim != 3 or arr.shape[1:] != (6, 6):
        raise ValueError(f"by_element values must be 6x6 matrices, got shape {arr.shape[1:]}")
    D_ref = arr[0].copy()
    piv = np.unravel_index(np.argmax(np.abs(D_ref)), D_ref.shape)
    if D_ref[piv] == 0:
        raise ValueError("by_element: the first element matrix is zero")
    s = arr[:, piv[0], piv[1]] / D_ref[piv]
    resid = np.abs(arr - s[:, None, None] * D_ref[None]).max(axis=(1, 2))
    scale = np.abs(arr).max(axis=(1, 2))
    if np.any(resid > 1e-12 * np.maximum(scale, 1e-300)):
        raise ValueError("by_element: element matrices are not scalar multiples of one another "
                         "(different Poisson ratios); the shared-K0 device operator needs D_e = s_e * D")
    return D_ref, s
