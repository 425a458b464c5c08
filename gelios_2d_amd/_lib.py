"""Loader of the HIP library libgelios.so (built in-tree by ``gelios_2d_amd.build``).

There is no CPU fallback: every product call goes through this library, and
:func:`lib` raises if it is missing or fails to load.
"""
from __future__ import annotations

import ctypes
import os

from .params import GelMcParams, GelMcStats, GelParams, GelTracking

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgelios.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i8p = ctypes.POINTER(ctypes.c_int8)


class GelGrid(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("x", _dp), ("y", _dp),
                ("cell_type", _i8p), ("outer_type", _i8p)]


_ip = ctypes.POINTER(ctypes.c_int)


class GelMesh(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int), ("n_cells", ctypes.c_int), ("x", _dp), ("y", _dp),
                ("point_type", _i8p), ("corner_offset", _ip), ("corners", _ip), ("cell_type", _i8p),
                ("face_offset", _ip), ("face_a", _ip), ("face_b", _ip), ("face_type", _ip), ("face_sosed", _ip),
                ("face_down", _ip), ("face_up", _ip)]


class GelMeshRails(ctypes.Structure):
    _fields_ = [("n_rails", ctypes.c_int), ("kind", _ip), ("m", _ip), ("s", _dp), ("all_offset", _ip),
                ("all", _ip), ("key", _ip), ("n_line", ctypes.c_int * 3), ("line", _ip), ("left", ctypes.c_double),
                ("d_chain", ctypes.c_int)]


class GelStats(ctypes.Structure):
    _fields_ = [("dt_last", ctypes.c_double), ("dt_next", ctypes.c_double), ("steps", ctypes.c_longlong),
                ("n_rho_clamp", ctypes.c_longlong), ("n_p_clamp", ctypes.c_longlong),
                ("n_numeric", ctypes.c_longlong)]


class GelError(RuntimeError):
    pass


# Every symbol include/gelios.h declares, with (restype, argtypes).
SIGNATURES = {
    "gel_params_reference": (None, [ctypes.POINTER(GelParams)]),
    "gel_ctx_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(GelGrid),
                                      ctypes.POINTER(GelParams)]),
    "gel_ctx_create_mesh": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(GelMesh),
                                           ctypes.POINTER(GelParams)]),
    "gel_ctx_destroy": (None, [ctypes.c_void_p]),
    "gel_upload_state": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "gel_download_state": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "gel_set_sources": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "gel_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(GelStats)]),
    "gel_step_moving": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.POINTER(GelStats)]),
    "gel_sync": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GelStats)]),
    "gel_set_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "gel_set_rows_per_wave": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "gel_enable_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "gel_kernel_timing": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "gel_flux_hllc2d_batch": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                             _dp, _dp]),
    "gel_set_source_table": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "gel_get_source_table": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "gel_get_source_flags": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "gel_set_motion": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_double]),
    "gel_tracking_reference": (None, [ctypes.POINTER(GelTracking)]),
    "gel_set_tracking": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GelTracking), _dp]),
    "gel_set_mesh_tracking": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GelTracking),
                                             ctypes.POINTER(GelMeshRails)]),
    "gel_get_nodes": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "gel_flux_hlld_batch": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp]),
    "gel_fp64_div_sqrt": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]),
    "gel_flux_hlld_batch_noalloc": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp]),
    "gel_mlp_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint64]),
    "gel_mlp_destroy": (None, [ctypes.c_void_p]),
    "gel_mlp_param_count": (ctypes.c_int, []),
    "gel_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "gel_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "gel_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _fp, _fp]),
    "gel_mlp_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _fp, _fp, _fp, _fp]),
    "gel_mlp_set_dataset": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _fp, _fp]),
    "gel_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, _fp]),
    "gel_mlp_set_scaling": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp, _fp, _fp]),
    "gel_mlp_last_error": (ctypes.c_char_p, []),
    "gel_mlp_set_output_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "gel_surrogate_features": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "gel_source_analytic": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "gel_surrogate_refresh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_longlong)]),
    "gel_flux_godunov_batch": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                              _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]),
    "gel_flux_godunov_wave": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                              _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]),
    "gel_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "gel_batch_destroy": (None, [ctypes.c_void_p]),
    "gel_batch_set_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "gel_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]),
    "gel_batch_prepare": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "gel_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), _dp]),
    "gel_batch_refresh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_longlong)]),
    "gel_batch_set_source_model": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_int]),
    "gel_batch_source_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                                              ctypes.POINTER(ctypes.c_longlong)]),
    "gel_mc_params_reference": (None, [ctypes.POINTER(GelMcParams)]),
    "gel_mc_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GelMcParams), _dp, ctypes.POINTER(GelMcStats)]),
    "gel_mc_rates": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp]),
    "gel_mc_exchange": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, _dp,
                                       ctypes.POINTER(ctypes.c_int)]),
    "gel_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "gel_last_error": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """Load libgelios.so; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GelError(f"{LIB_PATH} is not built; run gelios_2d_amd.build.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != 0:
        L = lib()
        raise GelError(f"{L.gel_strerror(status).decode()}: {L.gel_last_error().decode()}")


def dptr(a):
    return a.ctypes.data_as(_dp)


def i8ptr(a):
    return a.ctypes.data_as(_i8p)
