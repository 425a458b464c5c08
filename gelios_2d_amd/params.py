"""Sweep parameters: the ctypes mirror of ``gel_params`` (include/gelios.h).

The reference keeps these as compile-time macros (Help.h:39-64) and literals inside
Gran::Get_par (Gran.cpp:234-282); here they are run-time fields so that the
synthetic heliosphere workloads can set their own inflow while the parity runs use
the reference's values (:func:`reference_params`).
"""
from __future__ import annotations

import ctypes
import math

GEL_F_USUALY, GEL_F_INNER_SPHERE, GEL_F_EXTERN, GEL_F_INPUT, GEL_F_AXIS, GEL_F_UPPER_WALL = range(6)
GEL_C_1, GEL_C_2, GEL_C_3, GEL_C_4, GEL_C_5, GEL_C_CENTR, GEL_C_NO = range(7)
GEL_FLUX_HLL, GEL_FLUX_HLLC = 0, 1
GEL_PHYS_GAS, GEL_PHYS_MHD, GEL_PHYS_GAS4H, GEL_PHYS_GAS_INNER = 0, 1, 2, 3
NVAR_GAS4H = 21
NVAR_MHD = 9
GEL_SRC_NONE, GEL_SRC_MK1, GEL_SRC_TABLE, GEL_SRC_HYBRID = 0, 1, 2, 3
GEL_PREC_FP64, GEL_PREC_FP32_BULK = 0, 1
GEL_MK1_NFIELDS = 19
NVAR_GAS = 5

# Reference constants (Help.h).
RR_ = 390.222
CHI_REAL = 36.1275
VELOSITY_INF = -2.54338
A_2 = 0.1307345665
KN_ = 6.0
N_P_LISM = 3.0
N_H_LISM = 1.0
GA = 5.0 / 3.0


class GelParams(ctypes.Structure):
    """Field-for-field mirror of ``struct gel_params``."""

    _fields_ = [
        ("gamma", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("dt0", ctypes.c_double),
        ("r_skip", ctypes.c_double),
        ("flux", ctypes.c_int),
        ("physics", ctypes.c_int),
        ("source_mode", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("in_rho_div", ctypes.c_double),
        ("in_p_a", ctypes.c_double),
        ("in_p_den", ctypes.c_double),
        ("in_u", ctypes.c_double),
        ("in_q", ctypes.c_double),
        ("inf_rho", ctypes.c_double),
        ("inf_p", ctypes.c_double),
        ("inf_u", ctypes.c_double),
        ("inf_v", ctypes.c_double),
        ("inf_q", ctypes.c_double),
        ("extern_u", ctypes.c_double),
        ("src_coef", ctypes.c_double),
        ("sig_a", ctypes.c_double),
        ("sig_b", ctypes.c_double),
        ("inf_w", ctypes.c_double),
        ("inf_bx", ctypes.c_double),
        ("inf_by", ctypes.c_double),
        ("inf_bz", ctypes.c_double),
        ("src_coef_h", ctypes.c_double),
        ("h_pow", ctypes.c_double),
        ("h1_in_rho", ctypes.c_double),
        ("h1_in_p", ctypes.c_double),
        ("h4_inf_rho", ctypes.c_double),
        ("h4_inf_p", ctypes.c_double),
        ("h4_inf_u", ctypes.c_double),
        ("h4_inf_v", ctypes.c_double),
    ]

    def copy(self) -> "GelParams":
        p = GelParams()
        ctypes.pointer(p)[0] = self
        return p

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def reference_params() -> GelParams:
    """The reference's own constants: Help.h macros, Gran::Get_par literals, HLL flux
    as in Setka::Go_5_komponent__MK2 (``met = 0``, Setka.cpp:10211)."""
    p = GelParams()
    p.gamma = GA
    p.cfl = 0.1                       # kurant, Help.h:57
    p.dt0 = 0.00000001                # T[0] = T[1], Setka.cpp:10128-10129
    p.r_skip = 38.0 / RR_             # R11_, Help.h:81
    p.flux = GEL_FLUX_HLL
    p.physics = GEL_PHYS_GAS
    p.source_mode = GEL_SRC_NONE
    p.in_rho_div = CHI_REAL * CHI_REAL          # kv(chi_real), Gran.cpp:240
    p.in_p_a = CHI_REAL                         # ro*chi*chi/(ggg*10*10), Gran.cpp:241
    p.in_p_den = GA * 10.0 * 10.0
    p.in_u = CHI_REAL
    p.in_q = 1.0
    p.inf_rho, p.inf_p, p.inf_u, p.inf_v, p.inf_q = 1.0, 1.0, VELOSITY_INF, 0.0, 100.0
    p.extern_u = -0.1                           # Gran.cpp:216-219
    p.src_coef = N_P_LISM / KN_
    p.sig_a = 1.0
    p.sig_b = A_2
    p.src_coef_h = N_H_LISM / KN_                # Setka.cpp:9528-9547
    p.h_pow = 0.5                               # H_pow, Help.h:90
    p.h1_in_rho = 0.0000001                     # Gran.cpp:244-245
    p.h1_in_p = (0.0000001 * CHI_REAL * CHI_REAL / (GA * 14.1344 * 14.1344))
    p.h4_inf_rho, p.h4_inf_p, p.h4_inf_u, p.h4_inf_v = 1.0, 0.5, VELOSITY_INF, 0.0   # Gran.cpp:278-280
    return p


def heliosphere_params(*, flux: int = GEL_FLUX_HLLC, cfl: float = 0.4, rho_in: float = 1.0,
                       u_in: float = 10.0, p_in: float = 0.01, rho_inf: float = 0.01,
                       u_inf: float = 2.5, mach_inf: float = 2.0, gamma: float = GA) -> GelParams:
    """Synthetic single-fluid heliosphere: supersonic wind from the inner sphere r = 1
    (ro = rho_in/r^2, p = p_in/r^2 at r, |u| = u_in radial) against a uniform inflow
    along -x of density rho_inf, speed u_inf and Mach number mach_inf."""
    p = reference_params()
    p.gamma = gamma
    p.cfl = cfl
    p.flux = flux
    p.r_skip = 0.0
    p.in_rho_div = 1.0 / rho_in
    p.in_p_a = 1.0
    p.in_p_den = rho_in / p_in
    p.in_u = u_in
    p.in_q = 1.0
    p_inf = rho_inf * u_inf * u_inf / (gamma * mach_inf * mach_inf)
    p.inf_rho, p.inf_p, p.inf_u, p.inf_v, p.inf_q = rho_inf, p_inf, -u_inf, 0.0, 0.0
    return p



N_INNER = 20            # rail nodes of the inner block, Help.h:66
N_OUTER_SHOCK = 13      # Help.h:67


class GelTracking(ctypes.Structure):
    """Field-for-field mirror of ``struct gel_tracking`` (fitted-surface tracking)."""

    _fields_ = [
        ("row_inner_shock", ctypes.c_int),
        ("row_contact", ctypes.c_int),
        ("row_outer_shock", ctypes.c_int),
        ("outer_faces", ctypes.c_int),
        ("average_lines", ctypes.c_int),
        ("inner_rail_nodes", ctypes.c_int),
        ("r_inner", ctypes.c_double),
        ("r_mid", ctypes.c_double),
        ("r_outer", ctypes.c_double),
        ("p_inner", ctypes.c_double),
        ("p_mid", ctypes.c_double),
        ("p_outer", ctypes.c_double),
        ("s_outer", ctypes.c_double),
        ("koef", ctypes.c_double),
        ("k_contact", ctypes.c_double),
        ("k_inner", ctypes.c_double),
        ("k_outer", ctypes.c_double),
        ("tang_inner", ctypes.c_double),
        ("tang_outer", ctypes.c_double),
    ]

    def copy(self) -> "GelTracking":
        t = GelTracking()
        ctypes.pointer(t)[0] = self
        return t

    @property
    def rows(self):
        return self.row_inner_shock, self.row_contact, self.row_outer_shock


def reference_tracking(rows=(-1, -1, -1), outer_faces: int = 1 << 30) -> GelTracking:
    """The constants of Setka::Move_surface (Setka.cpp:3378-3844) and of the A rails of
    Setka::Move_Setka_Calculate_2 (Setka.cpp:4324-4381), with the Help.h radii.
    rows = node rows of the inner shock, the contact and the outer shock."""
    t = GelTracking()
    t.row_inner_shock, t.row_contact, t.row_outer_shock = rows
    t.outer_faces = outer_faces
    t.average_lines = 0
    t.inner_rail_nodes = N_INNER
    t.r_inner = 1.0 / RR_            # R1_
    t.r_mid = 50.0 / RR_             # R111_
    t.r_outer = 2200.0 / RR_         # R5_
    t.p_inner, t.p_mid, t.p_outer, t.s_outer = 1.3, 1.05, 1.76, 1.5708
    t.koef = 0.1
    t.k_contact, t.k_inner, t.k_outer = 0.3, 0.1, 1.0
    t.tang_inner, t.tang_outer = 0.1, 0.01
    return t


class GelMcParams(ctypes.Structure):
    """Field-for-field mirror of ``struct gel_mc_params`` (hydrogen Monte Carlo)."""

    _fields_ = [
        ("v_inf", ctypes.c_double),
        ("kn", ctypes.c_double),
        ("sig_a", ctypes.c_double),
        ("n_p_lism", ctypes.c_double),
        ("w_min", ctypes.c_double),
        ("n_primary", ctypes.c_longlong),
        ("seed", ctypes.c_ulonglong),
        ("max_segments", ctypes.c_int),
    ]


class GelMcStats(ctypes.Structure):
    """Field-for-field mirror of ``struct gel_mc_stats``."""

    _fields_ = [
        ("primaries_in", ctypes.c_longlong),
        ("atoms", ctypes.c_longlong),
        ("segments", ctypes.c_longlong),
        ("lost", ctypes.c_longlong),
        ("dropped", ctypes.c_longlong),
        ("generations", ctypes.c_int),
        ("ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


GEL_MC_NFIELDS = 29


def mc_reference_params(n_primary: int = 1_000_000, seed: int = 1, w_min: float = 0.25) -> GelMcParams:
    """The constants of the reference's Monte Carlo: Velosity_inf, Kn_, a_2, n_p_LISM_
    (Help.h:53-64) and weight_ (Help.h:93) as the secondaries' roulette weight."""
    p = GelMcParams()
    p.v_inf = -2.54338
    p.kn = 6.0
    p.sig_a = 0.1307345665
    p.n_p_lism = 3.0
    p.w_min = w_min
    p.n_primary = n_primary
    p.seed = seed
    p.max_segments = 1_000_000
    return p
