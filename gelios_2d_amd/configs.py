"""Synthetic heliosphere workloads (seeded) on polar grids.

Geometry: r log-spaced in [1, 400] (cells along each ray), polar angle in [0, pi]
(x = r cos(theta) along the symmetry axis, y = r sin(theta)).  Inner sphere r = 1:
supersonic wind ro = 1, |u| = 10 radial, p = 0.01.  Outer sphere: uniform inflow along
-x with ro = 0.01, |u| = 2.5, Mach 2, gamma = 5/3.  The initial state blends the
wind and the inflow with s(r) = (1 - tanh((r - r_b)/D))/2, r_b = 100, D = 20, and
carries a seeded +-1% density perturbation.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .grid import PolarGrid, polar_grid
from .params import GA, GEL_FLUX_HLLC, GEL_PHYS_MHD, GEL_PREC_FP64, GelParams, heliosphere_params


@dataclass
class Workload:
    name: str
    grid: PolarGrid
    params: GelParams
    prim0: np.ndarray
    steps: int
    motion: tuple = None          # (amp, period, phase) of the prescribed node motion
    src: np.ndarray = None        # [GEL_MK1_NFIELDS][ncells] hydrogen moments of the charge-exchange source


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


def hllc_gas_200x100() -> Workload:
    """Single-fluid HLLC, 200 x 100 cells, CFL 0.4, 100 steps, seed 1."""
    return heliosphere(200, 100, 100, seed=1, name="hllc-gas-200x100-static")


def large(ni: int = 4096, nj: int = 2048, steps: int = 20) -> Workload:
    """Same physics on a grid whose state (5 x 8 B x 2 buffers per cell) is far
    larger than the 256 MB Infinity Cache, for the HBM roofline."""
    return heliosphere(ni, nj, steps, seed=7, name=f"hllc-{ni}x{nj}")


B_ISM, B_ANGLE_DEG = 0.05, 60.0
MOTION_AMP, MOTION_PERIOD = 1e-3, 50


def motion_phase(seed: int) -> float:
    """Global phase of the prescribed motion, fixed by the seed."""
    return 2.0 * math.pi * ((seed * 0.6180339887498949) % 1.0)


def motion_scales(amp: float, period: int, phase: float) -> list:
    """s_k = 1 + amp sin(2 pi k/period + phase): the same expression, in the same
    order, as the library's motion table (glibc sin on the host)."""
    return [1.0 + amp * math.sin(2.0 * math.pi * k / period + phase) for k in range(period)]


def mhd_heliosphere(ni: int, nj: int, steps: int, seed: int, moving: bool = True, cfl: float = 0.4,
                    precision: int = GEL_PREC_FP64, b_ism: float = B_ISM, b_angle_deg: float = B_ANGLE_DEG,
                    name: str = "mhd-heliosphere") -> Workload:
    """Ideal MHD (HLLD), 9 variables, same geometry and flows as the gas workload;
    the interstellar inflow carries a field |B| = b_ism at b_angle_deg to the axis
    (in the meridional plane); the grid breathes radially by 1e-3 r with period 50
    steps (global phase fixed by the seed)."""
    w = heliosphere(ni, nj, steps, seed, cfl=cfl, name=name)
    prm = w.params
    prm.physics = GEL_PHYS_MHD
    prm.precision = precision
    a = math.radians(b_angle_deg)
    prm.inf_w, prm.inf_bx, prm.inf_by, prm.inf_bz = 0.0, b_ism * math.cos(a), b_ism * math.sin(a), 0.0
    ro, p, u, v, Q = w.prim0
    cx, cy = w.grid.cell_centers()
    r = np.hypot(cx, cy)
    s = 0.5 * (1.0 - np.tanh((r - R_BLEND) / D_BLEND))
    z = np.zeros_like(ro)
    q = np.stack([ro, p, u, v, z, (1 - s) * prm.inf_bx, (1 - s) * prm.inf_by, z, Q])
    motion = (MOTION_AMP, MOTION_PERIOD, motion_phase(seed)) if moving else None
    return Workload(name, w.grid, prm, np.ascontiguousarray(q), steps, motion)


def hlld_mhd_800x400(precision: int = GEL_PREC_FP64) -> Workload:
    """HLLD MHD, 800 x 400 cells, moving grid, 1000 steps, CFL 0.4, seed 2."""
    tag = "fp64" if precision == GEL_PREC_FP64 else "fp32bulk"
    return mhd_heliosphere(800, 400, 1000, seed=2, precision=precision, name=f"hlld-mhd-{tag}-800x400-moving")


def mhd_hydrogen_1600x800(precision: int = GEL_PREC_FP64, steps: int = 2000) -> Workload:
    """The multi-fluid form: HLLD MHD at 1600 x 800 cells on the moving grid, coupled to
    four frozen hydrogen populations (n_H = 0.1 (1 - exp(-r/50)), bulk velocity -2.5
    along the axis plus a +-5% seeded field, lognormal temperature around 1) by the
    charge-exchange source with sigma(U) = (1 - 0.07 ln U)^2 (Cell::Get_Sourse_MK1,
    Cell.cpp:498-665); inflow Mach 2, CFL 0.4, seed 3, 2000 steps.  The hydrogen fields
    are in .src."""
    from .surrogate import workload_with_hydrogen
    tag = "fp64" if precision == GEL_PREC_FP64 else "fp32bulk"
    w, f = workload_with_hydrogen(1600, 800, steps, INF_MACH, 0.1, 3, moving=True)
    w.params.precision = precision
    w.src = f
    w.name = f"hlld-mhd-{tag}-hydrogen-1600x800-moving"
    return w


def learned_mhd_800x400(precision: int = GEL_PREC_FP64, source_mode: int = None, steps: int = 1000) -> Workload:
    """The learned-source form of the 800 x 400 moving HLLD workload: the same grid,
    motion (seed 2), field and Mach-2 inflow, coupled to four frozen hydrogen populations
    (n_H = 0.1 (1 - exp(-r/50)), seeded perturbations) by the charge-exchange source with
    sigma(U) = (1 - 0.07 ln U)^2.  source_mode: GEL_SRC_HYBRID (default; the network's
    table, the analytic closure where its judge rejects it) or GEL_SRC_MK1 (the analytic
    closure every step, the plain form).  The hydrogen fields are in .src."""
    from .params import GEL_SRC_HYBRID
    from .surrogate import workload_with_hydrogen
    if source_mode is None:
        source_mode = GEL_SRC_HYBRID
    w, f = workload_with_hydrogen(800, 400, steps, INF_MACH, 0.1, 2, moving=True)
    w.params.precision = precision
    w.params.source_mode = source_mode
    w.src = f
    tag = "fp64" if precision == GEL_PREC_FP64 else "fp32bulk"
    kind = "learned" if source_mode == GEL_SRC_HYBRID else "analytic"
    w.name = f"hlld-mhd-{tag}-{kind}-source-800x400-moving"
    return w


def ensemble_cases(n_cases: int = 64, seed: int = 5, steps: int = 500, precision: int = None,
                   source_mode: int = None):
    """Batched ensemble: grid sizes log-uniform from 100 x 50 to 800 x 400, inflow Mach
    uniform in [1.5, 3], field angle uniform in [0, 90] deg, |B| uniform in [0.02, 0.1],
    hydrogen density uniform in [0.05, 0.2]; FP32-bulk HLLD with the learned source and
    its analytic fallback (GEL_SRC_HYBRID).  Returns a list of (workload, hydrogen
    fields)."""
    from .params import GEL_PREC_FP32_BULK, GEL_SRC_HYBRID
    from .surrogate import workload_with_hydrogen
    if precision is None:
        precision = GEL_PREC_FP32_BULK
    if source_mode is None:
        source_mode = GEL_SRC_HYBRID
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_cases):
        ni = int(round(math.exp(rng.uniform(math.log(100), math.log(800)))))
        nj = max(ni // 2, 3)
        mach = float(rng.uniform(1.5, 3.0))
        ang = float(rng.uniform(0.0, 90.0))
        bmag = float(rng.uniform(0.02, 0.1))
        nh = float(rng.uniform(0.05, 0.2))
        w, f = workload_with_hydrogen(ni, nj, steps, mach, nh, seed * 1000 + k, moving=False)
        a = math.radians(ang)
        w.params.inf_bx, w.params.inf_by = bmag * math.cos(a), bmag * math.sin(a)
        cx, cy = w.grid.cell_centers()
        s = 0.5 * (1.0 - np.tanh((np.hypot(cx, cy) - R_BLEND) / D_BLEND))
        w.prim0[5] = (1 - s) * w.params.inf_bx
        w.prim0[6] = (1 - s) * w.params.inf_by
        w.params.precision = precision
        w.params.source_mode = source_mode
        w.name = f"ensemble-case{k:02d}-{ni}x{nj}"
        out.append((w, f))
    return out


@dataclass
class TrackedCase:
    name: str
    grid: PolarGrid
    params: GelParams
    tracking: object              # params.GelTracking
    rails: np.ndarray             # nj+1 rail angles
    prim0: np.ndarray
    steps: int


def tracked_key_radii(s: float, theta1: float):
    """Inner shock, contact and outer shock radii (units of 390.222 AU, the reference's
    RR_) on the rail at angle s: about 98, 148 and 242 AU at the nose, widening to the
    flank."""
    q = (s / theta1) ** 2
    return 0.25 * (1.0 + 0.25 * q), 0.38 * (1.0 + 0.3 * q), 0.62 * (1.0 + 0.4 * q)


def tracked_heliosphere(nj: int, blocks, steps: int, seed: int, flux: int = 0,
                        theta1: float = 2.0 * math.pi / 3.0, outer_faces: int = None,
                        key_radii=None, name: str = "tracked-heliosphere") -> TrackedCase:
    """A heliosphere with fitted inner shock, contact and outer shock on the reference's
    A rails (Rail.cpp:19-75), the reference's own constants (Help.h) and boundary
    states: supersonic wind (the Inner_sphere formula, Gran.cpp:234-265) inside the
    inner shock (C_1 cells), the wind shocked by a strong shock between the inner shock
    and the contact, the inflow (Gran.cpp:276-282) shocked at Mach 1.97 between the
    contact and the outer shock, the undisturbed inflow outside; +-1% seeded density
    noise.  blocks = (M1, M2, M3, M4) rail nodes per block; key_radii(s) -> (inner shock,
    contact, outer shock) radii on the rail at angle s (default tracked_key_radii)."""
    from .grid import tracked_polar_grid
    from .params import GEL_C_1, GEL_C_2, reference_params, reference_tracking
    prm = reference_params()
    prm.flux = flux
    t = reference_tracking(outer_faces=outer_faces if outer_faces is not None else 1 << 30)
    if key_radii is None:
        key_radii = lambda s: tracked_key_radii(s, theta1)   # noqa: E731
    g, rails, rows = tracked_polar_grid(nj, blocks, key_radii, t, theta1)
    t.row_inner_shock, t.row_contact, t.row_outer_shock = rows
    gam = prm.gamma
    cx, cy = g.cell_centers()
    r = np.hypot(cx, cy)
    i = np.repeat(np.arange(g.ni), g.nj)
    ro_w = 1.0 / (prm.in_rho_div * r * r)
    p_w = ro_w * prm.in_p_a * prm.in_p_a / prm.in_p_den
    u_w, v_w = prm.in_u * cx / r, prm.in_u * cy / r
    M = abs(prm.inf_u) / math.sqrt(gam * prm.inf_p / prm.inf_rho)
    comp = (gam + 1) * M * M / ((gam - 1) * M * M + 2)
    p_ratio = (2 * gam * M * M - (gam - 1)) / (gam + 1)
    ro = np.where(i < rows[0], ro_w, 0.0)
    ro = np.where((i >= rows[0]) & (i < rows[1]), 4.0 * ro_w, ro)
    ro = np.where((i >= rows[1]) & (i < rows[2]), prm.inf_rho * comp, ro)
    ro = np.where(i >= rows[2], prm.inf_rho, ro)
    p = np.where(i < rows[0], p_w, 0.0)
    p = np.where((i >= rows[0]) & (i < rows[1]), 0.75 * ro_w * prm.in_u ** 2, p)
    p = np.where((i >= rows[1]) & (i < rows[2]), prm.inf_p * p_ratio, p)
    p = np.where(i >= rows[2], prm.inf_p, p)
    u = np.where(i < rows[0], u_w, np.where(i < rows[1], 0.25 * u_w,
                 np.where(i < rows[2], prm.inf_u / comp, prm.inf_u)))
    v = np.where(i < rows[0], v_w, np.where(i < rows[1], 0.25 * v_w, 0.0))
    Q = np.where(i < rows[1], ro, prm.inf_q)
    rng = np.random.default_rng(seed)
    ro = ro * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, ro.shape))
    g.cell_type = np.where(i < rows[0], GEL_C_1, GEL_C_2).astype(np.int8)
    prim0 = np.ascontiguousarray(np.stack([ro, p, u, v, Q]))
    return TrackedCase(name, g, prm, t, rails, prim0, steps)


def tracked_spherical_ts(s: float, theta1: float = 2.0 * math.pi / 3.0):
    """Key radii with a spherical inner shock at 0.25 (contact and outer shock as in
    tracked_key_radii)."""
    _, hp, bs = tracked_key_radii(s, theta1)
    return 0.25, hp, bs


def multifluid_heliosphere(ni: int, nj: int, steps: int, seed: int, flux: int = 1,
                           name: str = "plasma-4h-heliosphere") -> Workload:
    """The plasma and the four hydrogen fluids of Setka::Go_5_komponent_2 on a polar
    grid r in [4, 12] (units of 390.222 AU), theta in [0, pi], with the reference's
    constants: the supersonic wind (the Inner_sphere formula, Gran.cpp:234-265) blended
    into the inflow (Gran.cpp:276-282) at r = 7; cell types by radius select the
    hydrogen population each region feeds (C_1 below 5, C_2 to 7, C_3 to 8, C_4 to 10,
    C_5 beyond; Setka.cpp:9554-9586); H4 starts as the interstellar hydrogen (1, 0.5,
    -2.54338, 0), H1..H3 as thinner, slower copies; +-1% seeded density noise."""
    from .grid import polar_grid
    from .params import (GEL_C_1, GEL_C_2, GEL_C_3, GEL_C_4, GEL_C_5, GEL_PHYS_GAS4H,
                         reference_params)
    prm = reference_params()
    prm.physics = GEL_PHYS_GAS4H
    prm.flux = flux
    g = polar_grid(ni, nj, 4.0, 12.0, spacing="log")
    cx, cy = g.cell_centers()
    r = np.hypot(cx, cy)
    t = np.full(g.ncells, GEL_C_5, np.int8)
    for lim, ty in ((10.0, GEL_C_4), (8.0, GEL_C_3), (7.0, GEL_C_2), (5.0, GEL_C_1)):
        t[r < lim] = ty
    g.cell_type = t
    ro_w = 1.0 / (prm.in_rho_div * r * r)
    p_w = ro_w * prm.in_p_a * prm.in_p_a / prm.in_p_den
    s = 0.5 * (1 - np.tanh((r - 7.0) / 1.5))
    ro = s * ro_w + (1 - s) * prm.inf_rho
    p = s * p_w + (1 - s) * prm.inf_p
    u = s * prm.in_u * cx / r + (1 - s) * prm.inf_u
    v = s * prm.in_u * cy / r + (1 - s) * prm.inf_v
    Q = s * ro_w + (1 - s) * prm.inf_q
    rng = np.random.default_rng(seed)
    fields = [ro * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, ro.shape)), p, u, v, Q]
    for k, scale in enumerate((0.05, 0.1, 0.2, 1.0)):
        rh = scale * prm.h4_inf_rho * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, ro.shape))
        fields += [rh, scale * prm.h4_inf_p + 0.0 * r, (0.5 + 0.5 * scale) * prm.h4_inf_u + 0.0 * r, 0.0 * r]
    return Workload(name, g, prm, np.ascontiguousarray(np.stack(fields)), steps)


def inner_region(ni: int, nj: int, steps: int, seed: int, flux: int = 1, sources: bool = False,
                 name: str = "inner-region") -> Workload:
    """The region the inner-region solver (Setka::Go_stationary_5_komponent_inner_MK2)
    advances: r in [R1_, R111_] = [1, 50] AU / 390.222, theta in [0, pi], log-spaced
    rows; the supersonic wind of the Inner_sphere formula (Gran.cpp:234-265) with +-1%
    seeded density noise; C_1 cells up to 3/4 of the rows, C_2 beyond, C_centr cells at
    both axis ends of the first row; the outer faces copy the cell (Upper_wall).  With
    sources: seeded hydrogen moments for the MK1 closure."""
    from .grid import polar_grid
    from .params import (GEL_C_1, GEL_C_2, GEL_C_CENTR, GEL_F_UPPER_WALL, GEL_PHYS_GAS_INNER,
                         GEL_SRC_MK1, RR_, reference_params)
    prm = reference_params()
    prm.physics = GEL_PHYS_GAS_INNER
    prm.flux = flux
    if sources:
        prm.source_mode = GEL_SRC_MK1
    g = polar_grid(ni, nj, 1.0 / RR_, 50.0 / RR_, spacing="log")
    i = np.repeat(np.arange(ni), nj)
    t = np.where(i < (3 * ni) // 4, GEL_C_1, GEL_C_2).astype(np.int8)
    t[0] = GEL_C_CENTR
    t[nj - 1] = GEL_C_CENTR
    g.cell_type = t
    g.outer_type = np.full(nj, GEL_F_UPPER_WALL, np.int8)
    cx, cy = g.cell_centers()
    r = np.hypot(cx, cy)
    ro = 1.0 / (prm.in_rho_div * r * r)
    p = ro * prm.in_p_a * prm.in_p_a / prm.in_p_den
    u, v = prm.in_u * cx / r, prm.in_u * cy / r
    rng = np.random.default_rng(seed)
    prim = np.stack([ro * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, ro.shape)), p, u, v, ro.copy()])
    w = Workload(name, g, prm, np.ascontiguousarray(prim), steps)
    if sources:
        nc = g.ncells
        w.src = np.ascontiguousarray(np.concatenate([rng.uniform(0.05, 0.5, (4, nc)), rng.normal(-2.0, 0.5, (4, nc)),
                                                     rng.normal(0.0, 0.5, (4, nc)), rng.uniform(0.5, 2.0, (4, nc)),
                                                     rng.uniform(0.5, 2.0, (3, nc))]))
    return w
