"""Synthetic heliosphere workloads (seeded) on polar grids.

Geometry: r log-spaced in [1, 400] (cells along each ray), polar angle in [0, pi]
(x = r cos(theta) along the symmetry axis, y = r sin(theta)).  Inner sphere r = 1:
supersonic wind ro = 1, |u| = 10 radial, p = 0.01.  Outer sphere: uniform inflow along
-x with ro = 0.01, |u| = 2.5, Mach 2, gamma = 5/3.  The initial state blends the
wind and the inflow with s(r) = (1 - tanh((r - r_b)/D))/2, r_b = 100, D = 20, and
carries a seeded +-1% density perturbation.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .grid import PolarGrid, polar_grid
from .params import GA, GEL_FLUX_HLLC, GelParams, heliosphere_params


@dataclass
class Workload:
    name: str
    grid: PolarGrid
    params: GelParams
    prim0: np.ndarray
    steps: int


WIND_RHO, WIND_U, WIND_P = 1.0, 10.0, 0.01
INF_RHO, INF_U, INF_MACH = 0.01, 2.5, 2.0
R_IN, R_OUT = 1.0, 400.0
R_BLEND, D_BLEND = 100.0, 20.0


def blend_state(g: PolarGrid, seed: int, gamma: float = GA) -> np.ndarray:
    cx, cy = g.cell_centers()
    r = np.hypot(cx, cy)
    ro_w = WIND_RHO / r ** 2
    p_w = WIND_P * r ** (-2.0 * gamma)
    u_w, v_w = WIND_U * cx / r, WIND_U * cy / r
    p_inf = INF_RHO * INF_U ** 2 / (gamma * INF_MACH ** 2)
    s = 0.5 * (1.0 - np.tanh((r - R_BLEND) / D_BLEND))
    ro = s * ro_w + (1 - s) * INF_RHO
    p = s * p_w + (1 - s) * p_inf
    u = s * u_w + (1 - s) * (-INF_U)
    v = s * v_w
    Q = s * ro_w
    rng = np.random.default_rng(seed)
    ro = ro * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, ro.shape))
    return np.ascontiguousarray(np.stack([ro, p, u, v, Q]))


def heliosphere(ni: int, nj: int, steps: int, seed: int, flux: int = GEL_FLUX_HLLC,
                cfl: float = 0.4, name: str = "heliosphere") -> Workload:
    g = polar_grid(ni, nj, R_IN, R_OUT, spacing="log")
    prm = heliosphere_params(flux=flux, cfl=cfl, rho_in=WIND_RHO, u_in=WIND_U, p_in=WIND_P,
                             rho_inf=INF_RHO, u_inf=INF_U, mach_inf=INF_MACH)
    return Workload(name, g, prm, blend_state(g, seed), steps)


def config1() -> Workload:
    """Single-fluid HLLC, 200 x 100 cells, CFL 0.4, 100 steps, seed 1."""
    return heliosphere(200, 100, 100, seed=1, name="cfg1-hllc-200x100")


def large(ni: int = 4096, nj: int = 2048, steps: int = 20) -> Workload:
    """Same physics on a grid whose state (5 x 8 B x 2 buffers per cell) is far
    larger than the 256 MB Infinity Cache, for the HBM roofline."""
    return heliosphere(ni, nj, steps, seed=7, name=f"hllc-{ni}x{nj}")
