"""Many independent heliosphere cases advanced by one persistent launch per step.

Each case is a :class:`~gelios_2d_amd.setka.Setka` (its own grid, state, inflow,
sources and CFL time step); the batch owns the launch (csrc/gelios_sweep.hip,
sweep_batch_kernel), replays whole 6-step cycles from a captured HIP graph and
chooses its launch shape with the project autotuner (gel_batch_autotune).
"""
from __future__ import annotations

import ctypes

from . import _lib


class Batch:
    def __init__(self, cases):
        self.cases = list(cases)
        arr = (ctypes.c_void_p * len(self.cases))(*[c._ctx.value for c in self.cases])
        self._h = ctypes.c_void_p()
        _lib.check(_lib.lib().gel_batch_create(ctypes.byref(self._h), len(self.cases), arr))

    def close(self):
        if self._h:
            _lib.lib().gel_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ncells(self) -> int:
        return sum(c.ncells for c in self.cases)

    def set_launch(self, rows: int = 0, workgroups: int = 0) -> None:
        _lib.check(_lib.lib().gel_batch_set_launch(self._h, int(rows), int(workgroups)))

    def step(self, nsteps: int, graph: bool = True) -> float:
        """Advance every case; returns the device time in milliseconds."""
        ms = ctypes.c_double()
        _lib.check(_lib.lib().gel_batch_step(self._h, int(nsteps), 1 if graph else 0, ctypes.byref(ms)))
        return ms.value

    def autotune(self, trial_steps: int = 6):
        r, w, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _lib.check(_lib.lib().gel_batch_autotune(self._h, int(trial_steps), ctypes.byref(r), ctypes.byref(w),
                                                 ctypes.byref(ms)))
        return r.value, w.value, ms.value
