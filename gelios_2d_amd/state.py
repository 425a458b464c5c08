"""Primitive <-> conserved state (variable-major SoA arrays).

Gas primitives (ro, p, u, v, Q) -> conserved (ro, ro u, ro v, E, Q) with
E = p/(gamma-1) + ro (u^2+v^2)/2, as the update forms it (Setka.cpp:10444-10446).
MHD primitives (ro, p, u, v, w, bx, by, bz, Q) -> (ro, ro u, ro v, ro w, E, bx, by, bz, Q)
with E = p/(gamma-1) + ro |u|^2/2 + |B|^2/(8 pi) (Gaussian B).
"""
from __future__ import annotations

import numpy as np


def to_conserved(prim, gamma: float = 5.0 / 3.0) -> np.ndarray:
    q = np.asarray(prim, np.float64)
    if q.shape[0] == 5:
        ro, p, u, v, Q = q
        return np.stack([ro, ro * u, ro * v, p / (gamma - 1) + ro * (u * u + v * v) * 0.5, Q])
    ro, p, u, v, w, bx, by, bz, Q = q
    B2 = bx * bx + by * by + bz * bz
    E = p / (gamma - 1) + ro * (u * u + v * v + w * w) * 0.5 + B2 / (8 * np.pi)
    return np.stack([ro, ro * u, ro * v, ro * w, E, bx, by, bz, Q])


def rel_err_conserved(a, b, gamma: float = 5.0 / 3.0) -> float:
    """max over conserved variables of max|a-b| / max|b|."""
    A, B = to_conserved(a, gamma), to_conserved(b, gamma)
    scale = np.maximum(np.abs(B).max(axis=1), 1e-300)
    return float(np.max(np.abs(A - B).max(axis=1) / scale))


def l1_rel(a, b, k: int) -> float:
    """L1 error of variable k relative to the L1 norm of b's variable k."""
    a = np.asarray(a)[k]
    b = np.asarray(b)[k]
    return float(np.sum(np.abs(a - b)) / max(np.sum(np.abs(b)), 1e-300))


def linf_rel(a, b, k: int) -> float:
    """Largest error of variable k relative to the largest magnitude of b's variable k."""
    a = np.asarray(a)[k]
    b = np.asarray(b)[k]
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
