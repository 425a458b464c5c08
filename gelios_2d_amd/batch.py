"""Many independent heliosphere cases advanced by one persistent launch per step.

Each case is a :class:`~gelios_2d_amd.setka.Setka` (its own grid, state, inflow,
sources and CFL time step); the batch owns the launch (csrc/gelios_sweep.hip,
sweep_batch_kernel), replays whole cycles of steps from a captured HIP graph,
chooses its launch shape with the project autotuner (gel_batch_autotune) and, with
a source model attached, refreshes every case's learned charge-exchange source in
one fused launch every few steps (gel_batch_set_source_model).
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

    def prepare(self, nsteps: int) -> None:
        """Capture the graphs that step(nsteps) replays next, running nothing."""
        _lib.check(_lib.lib().gel_batch_prepare(self._h, int(nsteps)))

    def autotune(self, trial_steps: int = 6):
        r, w, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _lib.check(_lib.lib().gel_batch_autotune(self._h, int(trial_steps), ctypes.byref(r), ctypes.byref(w),
                                                 ctypes.byref(ms)))
        return r.value, w.value, ms.value

    def refresh(self, model, tol: float = 1e-3, judge=True) -> int:
        """Refresh every case's source table from its current state now; returns the
        number of cells that fell back to the analytic source.  judge: False / 0 none,
        True / 1 every cell, 2 (GEL_SRC_HYBRID) only the cells not flagged already."""
        n = ctypes.c_longlong()
        _lib.check(_lib.lib().gel_batch_refresh(self._h, model.mlp._h, float(tol), int(judge),
                                                ctypes.byref(n)))
        return n.value

    def set_source_model(self, model, every: int, tol: float = 1e-3, judge=True) -> None:
        """Refresh the source tables inside step() after every `every`-th step the
        batch advances from now on, over all calls whatever their lengths (0: off);
        judge as in refresh()."""
        self._model = model          # the batch holds the network's handle
        _lib.check(_lib.lib().gel_batch_set_source_model(self._h, model.mlp._h if model else None, int(every),
                                                         float(tol), int(judge)))

    def source_stats(self):
        """(refreshes, fallback cells) since the last call."""
        r, n = ctypes.c_longlong(), ctypes.c_longlong()
        _lib.check(_lib.lib().gel_batch_source_stats(self._h, ctypes.byref(r), ctypes.byref(n)))
        return r.value, n.value
