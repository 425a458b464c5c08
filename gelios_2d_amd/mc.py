"""The Monte Carlo of the interstellar hydrogen on the device (gel_mc_run): the
trajectory tracer that gives the sweep its charge-exchange inputs, in place of the
reference's Setka::MK_start_new (Setka.cpp:11989-12535) with Setka::Fly_exchenge_Imit_Korol
(:12814-13981) and the moments' post-processing (:12241-12449).

    from gelios_2d_amd.mc import run_mc
    fields, stats = run_mc(setka, n_primary=2_000_000)   # 29 x ncells, see FIELDS
    # a GEL_SRC_MK1 context now sweeps with fields 0..18 as its source inputs
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .params import GEL_MC_NFIELDS, GelMcParams, GelMcStats, mc_reference_params

FIELDS = ([f"H_n{k}" for k in range(1, 5)] + [f"H_u{k}" for k in range(1, 5)] + [f"H_v{k}" for k in range(1, 5)] +
          [f"H_T{k}" for k in range(1, 5)] + ["k_u", "k_v", "k_T", "I_u", "I_v", "I_T", "II_u", "II_v", "II_T",
                                               "F_n", "F_u", "F_v", "F_T"])


def run_mc(setka, params: GelMcParams | None = None, **overrides):
    """Run the Monte Carlo on a context's grid and plasma; returns (fields [29][ncells],
    stats dict).  overrides set fields of the parameters (n_primary, seed, w_min, ...)."""
    p = params if params is not None else mc_reference_params()
    for k, v in overrides.items():
        setattr(p, k, v)
    out = np.zeros((GEL_MC_NFIELDS, setka.ncells))
    st = GelMcStats()
    _lib.check(_lib.lib().gel_mc_run(setka._ctx, ctypes.byref(p), _lib.dptr(out), ctypes.byref(st)))
    return out, st.as_dict()


def rates(u, cp):
    """Device Velosity_1..3 and int_1..3 at (u, cp): [n][6]."""
    u = np.ascontiguousarray(u, np.float64)
    cp = np.ascontiguousarray(np.broadcast_to(cp, u.shape), np.float64)
    out = np.zeros((len(u), 6))
    _lib.check(_lib.lib().gel_mc_rates(len(u), _lib.dptr(u), _lib.dptr(cp), _lib.dptr(out)))
    return out


def exchange(inp, seeds, reps):
    """Device exchange-velocity sampler from seeded sensors: (out [n][reps][4], state [n][3])."""
    inp = np.ascontiguousarray(inp, np.float64)
    seeds = np.ascontiguousarray(seeds, np.int32)
    n = len(inp)
    out = np.zeros((n, reps, 4))
    state = np.zeros((n, 3), np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    _lib.check(_lib.lib().gel_mc_exchange(n, _lib.dptr(inp), seeds.ctypes.data_as(ip), int(reps), _lib.dptr(out),
                                          state.ctypes.data_as(ip)))
    return out, state
