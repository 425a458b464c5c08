"""The learned charge-exchange source: a 14 -> 4 x 128 SiLU -> 5 MLP on MFMA.

Forward, backward and Adam run in libgelios.so (csrc/surrogate.hip): BF16 weights
and activations, FP32 accumulation.  This module is the host driver: parameter I/O,
training on the device, and the feature/target scaling of the source model.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

N_IN, N_OUT = 14, 5
_fp = ctypes.POINTER(ctypes.c_float)


def _f(a):
    return a.ctypes.data_as(_fp)


def _setup(L):
    if getattr(L, "_mlp_ready", False):
        return L
    L.gel_mlp_create.restype = ctypes.c_int
    L.gel_mlp_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint64]
    L.gel_mlp_destroy.restype = None
    L.gel_mlp_destroy.argtypes = [ctypes.c_void_p]
    L.gel_mlp_param_count.restype = ctypes.c_int
    L.gel_mlp_param_count.argtypes = []
    for name in ("gel_mlp_get_params", "gel_mlp_set_params"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [ctypes.c_void_p, _fp]
    L.gel_mlp_forward.restype = ctypes.c_int
    L.gel_mlp_forward.argtypes = [ctypes.c_void_p, ctypes.c_int, _fp, _fp]
    L.gel_mlp_grad.restype = ctypes.c_int
    L.gel_mlp_grad.argtypes = [ctypes.c_void_p, ctypes.c_int, _fp, _fp, _fp, _fp]
    L.gel_mlp_set_dataset.restype = ctypes.c_int
    L.gel_mlp_set_dataset.argtypes = [ctypes.c_void_p, ctypes.c_int, _fp, _fp]
    L.gel_mlp_train.restype = ctypes.c_int
    L.gel_mlp_train.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, _fp]
    L.gel_mlp_last_error.restype = ctypes.c_char_p
    L._mlp_ready = True
    return L


def _check(st):
    if st != 0:
        L = _lib.lib()
        raise _lib.GelError(f"{L.gel_strerror(st).decode()}: {L.gel_mlp_last_error().decode()}")


class MLP:
    """Device MLP with FP32 master parameters and BF16 compute copies."""

    def __init__(self, seed: int = 0):
        self._L = _setup(_lib.lib())
        self._h = ctypes.c_void_p()
        _check(self._L.gel_mlp_create(ctypes.byref(self._h), ctypes.c_uint64(seed)))
        self.n_params = self._L.gel_mlp_param_count()

    def close(self):
        if self._h:
            self._L.gel_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        _check(self._L.gel_mlp_get_params(self._h, _f(p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(self._L.gel_mlp_set_params(self._h, _f(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_IN)
        y = np.empty((x.shape[0], N_OUT), np.float32)
        _check(self._L.gel_mlp_forward(self._h, x.shape[0], _f(x), _f(y)))
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_IN)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, N_OUT)
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_float()
        _check(self._L.gel_mlp_grad(self._h, x.shape[0], _f(x), _f(t), _f(g), ctypes.byref(loss)))
        return loss.value, g

    def set_dataset(self, x, t) -> None:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_IN)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, N_OUT)
        _check(self._L.gel_mlp_set_dataset(self._h, x.shape[0], _f(x), _f(t)))

    def train(self, steps: int, batch: int, lr: float) -> np.ndarray:
        loss = np.empty(max(steps, 1), np.float32)
        _check(self._L.gel_mlp_train(self._h, int(steps), int(batch), float(lr), _f(loss)))
        return loss[:steps]
