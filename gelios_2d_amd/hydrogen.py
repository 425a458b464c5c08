"""Frozen interstellar hydrogen for the charge-exchange source (synthetic).

Four populations, as the reference carries them (Parametr::H_n/H_u/H_v/H_T[4],
Cell.h:55-58), with analytic seeded profiles:
  n_k(r) = n0 (1 - exp(-r/50)) for every population,
  bulk velocity -u_inf along the axis plus a +-5% seeded random field,
  temperature lognormal around 1.
Populations 1-2 share one velocity/temperature field and 3-4 another, so that the
source is an exact function of the two group moments the surrogate sees.
The k factors (Monte Carlo / closure ratios, Setka.cpp:12241-12449) are 1.
"""
from __future__ import annotations

import numpy as np

from .grid import PolarGrid
from .params import GEL_MK1_NFIELDS


def populations(g: PolarGrid, n0: float = 0.1, u_inf: float = 2.5, seed: int = 0,
                frac=(0.5, 0.2, 0.2, 0.1), t_sigma: float = 0.2) -> np.ndarray:
    """[GEL_MK1_NFIELDS][ncells] fields: H_n[4], H_u[4], H_v[4], H_T[4], k_u, k_v, k_T."""
    rng = np.random.default_rng(seed)
    cx, cy = g.cell_centers()
    r = np.hypot(cx, cy)
    nc = g.ncells
    base = n0 * (1.0 - np.exp(-r / 50.0))
    f = np.zeros((GEL_MK1_NFIELDS, nc))
    for grp in range(2):
        du = 0.05 * u_inf * rng.uniform(-1, 1, (2, nc))
        T = np.exp(t_sigma * rng.standard_normal(nc))
        for k in (2 * grp, 2 * grp + 1):
            f[k] = 4.0 * frac[k] * base          # fractions sum to 1 over the 4 populations x 4
            f[4 + k] = -u_inf + du[0]
            f[8 + k] = du[1]
            f[12 + k] = T
    f[16:19] = 1.0
    return np.ascontiguousarray(f)
