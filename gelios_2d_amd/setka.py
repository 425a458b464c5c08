"""Host-side mirror of the reference's sweep interface, over the C-ABI.

The reference drives its sweep through the C++ object ``Setka`` (Setka.h:21-235):
cells carry ``par[2]`` double buffers (Cell.h:98), ``Go_5_komponent__MK2(step)``
advances ``step`` steps (Setka.h:186, Setka.cpp:10121-10482) and
``HLLC_2d_Korolkov_b_s`` evaluates one face flux (Setka.h:187-189).  :class:`Setka`
keeps those names; the state lives in device memory of one MI355X and every step
runs in libgelios.so.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .grid import MAXF, PolarGrid
from .params import GEL_MK1_NFIELDS, GEL_PHYS_GAS4H, GEL_PHYS_MHD, NVAR_GAS, NVAR_GAS4H, NVAR_MHD, GelParams


class Setka:
    """A polar grid, its state on the GPU, and the sweep.

    ``prim`` arrays are variable-major (ro, p, u, v, Q) x ncells, cell (i, j) at
    ``i*nj + j`` -- the fields of Parametr (Cell.h:21-25) as structure of arrays.
    """

    def __init__(self, grid: PolarGrid, params: GelParams):
        self.grid = grid
        self.params = params.copy()
        self._ctx = ctypes.c_void_p()
        g = _lib.GelGrid(grid.ni, grid.nj, _lib.dptr(grid.x), _lib.dptr(grid.y),
                         _lib.i8ptr(grid.cell_type), _lib.i8ptr(grid.outer_type))
        _lib.check(_lib.lib().gel_ctx_create(ctypes.byref(self._ctx), ctypes.byref(g), ctypes.byref(self.params)))
        self.stats = _lib.GelStats()
        self.nvar = {GEL_PHYS_MHD: NVAR_MHD, GEL_PHYS_GAS4H: NVAR_GAS4H}.get(self.params.physics, NVAR_GAS)

    def close(self):
        if self._ctx:
            _lib.lib().gel_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @classmethod
    def from_mesh(cls, mesh, params: GelParams) -> "Setka":
        """A context on a general mesh (:class:`~gelios_2d_amd.mesh.Mesh`: the
        reference's own block grid, or one read from its checkpoint): the FP64 gas
        sweep, static (Go_5_komponent__MK2) or with next node positions (step_moving)."""
        self = cls.__new__(cls)
        self.grid = mesh
        self.params = params.copy()
        self._ctx = ctypes.c_void_p()
        nc = mesh.ncells
        off = np.zeros(nc + 1, np.int32)
        off[1:] = np.cumsum(mesh.nface)
        keep = dict(x=np.ascontiguousarray(mesh.x, np.float64), y=np.ascontiguousarray(mesh.y, np.float64),
                    pt=np.ascontiguousarray(mesh.ptype, np.int8), coff=np.ascontiguousarray(mesh.coff, np.int32),
                    corners=np.ascontiguousarray(mesh.corners, np.int32),
                    ct=np.ascontiguousarray(mesh.cell_type, np.int8), foff=off)
        sel = np.concatenate([np.arange(c * MAXF, c * MAXF + mesh.nface[c]) for c in range(nc)])
        for k in ("fa", "fb", "ftype", "fnbr", "fdown", "fup"):
            keep[k] = np.ascontiguousarray(getattr(mesh, k)[sel], np.int32)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))   # noqa: E731
        g = _lib.GelMesh(mesh.nnodes, nc, _lib.dptr(keep["x"]), _lib.dptr(keep["y"]), _lib.i8ptr(keep["pt"]),
                         ip(keep["coff"]), ip(keep["corners"]), _lib.i8ptr(keep["ct"]), ip(keep["foff"]),
                         *[ip(keep[k]) for k in ("fa", "fb", "ftype", "fnbr", "fdown", "fup")])
        _lib.check(_lib.lib().gel_ctx_create_mesh(ctypes.byref(self._ctx), ctypes.byref(g), ctypes.byref(self.params)))
        self.stats = _lib.GelStats()
        self.nvar = NVAR_GAS
        return self

    @property
    def ncells(self) -> int:
        return self.grid.ncells

    # -- state ------------------------------------------------------------------
    def upload(self, prim: np.ndarray) -> None:
        a = np.ascontiguousarray(prim, np.float64).reshape(self.nvar, self.ncells)
        _lib.check(_lib.lib().gel_upload_state(self._ctx, _lib.dptr(a)))

    def download(self) -> np.ndarray:
        out = np.empty((self.nvar, self.ncells), np.float64)
        _lib.check(_lib.lib().gel_download_state(self._ctx, _lib.dptr(out)))
        return out

    def set_sources(self, fields: np.ndarray) -> None:
        a = np.ascontiguousarray(fields, np.float64).reshape(GEL_MK1_NFIELDS, self.ncells)
        _lib.check(_lib.lib().gel_set_sources(self._ctx, _lib.dptr(a)))

    def set_motion(self, amp: float, period: int, phase: float) -> None:
        """Prescribed breathing of the grid: nodes at base * (1 + amp sin(2 pi k/period
        + phase)), k = step mod period; steps become moving steps."""
        _lib.check(_lib.lib().gel_set_motion(self._ctx, float(amp), int(period), float(phase)))

    def set_tracking(self, tracking, rails) -> None:
        """Fitted-surface tracking (Setka::Move_surface, Setka.cpp:3378-3844, and the A
        rails of Setka::Move_Setka_Calculate_2, Setka.cpp:4324-4381): every step moves
        the three surface rows of ``tracking`` (a :class:`~.params.GelTracking`) and
        redistributes the rails (node columns at angles ``rails``) before the sweep.
        ``None`` turns it off."""
        if tracking is None:
            _lib.check(_lib.lib().gel_set_tracking(self._ctx, None, None))
            return
        r = np.ascontiguousarray(rails, np.float64)
        if r.shape != (self.grid.nj + 1,):
            raise ValueError("one rail angle per node column")
        _lib.check(_lib.lib().gel_set_tracking(self._ctx, ctypes.byref(tracking), _lib.dptr(r)))

    def set_mesh_tracking(self, tracking, rails) -> None:
        """Fitted-surface tracking on a general-mesh context (the reference's own block
        grid): every step runs Setka::Move_surface over the three lines and
        Setka::Move_Setka_Calculate_2 over the four rail families (Setka.cpp:3378-3844,
        :4324-4759) on the device before the sweep.  ``rails`` is a
        :class:`~.mesh.MeshRails`; ``None`` tracking turns it off."""
        if tracking is None:
            _lib.check(_lib.lib().gel_set_mesh_tracking(self._ctx, None, None))
            return
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))   # noqa: E731
        keep = dict(kind=np.ascontiguousarray(rails.kind, np.int32), m=np.ascontiguousarray(rails.m, np.int32),
                    s=np.ascontiguousarray(rails.s, np.float64),
                    off=np.ascontiguousarray(rails.all_offset, np.int32), all=np.ascontiguousarray(rails.all, np.int32),
                    key=np.ascontiguousarray(rails.key, np.int32),
                    line=np.ascontiguousarray(np.concatenate([np.asarray(l, np.int32).reshape(-1, 4)
                                                              for l in rails.lines]), np.int32))
        r = _lib.GelMeshRails(len(keep["kind"]), ip(keep["kind"]), ip(keep["m"]), _lib.dptr(keep["s"]), ip(keep["off"]),
                              ip(keep["all"]), ip(keep["key"]),
                              (ctypes.c_int * 3)(*[len(np.asarray(l).reshape(-1, 4)) for l in rails.lines]),
                              ip(keep["line"]), float(rails.left), int(rails.d_chain))
        _lib.check(_lib.lib().gel_set_mesh_tracking(self._ctx, ctypes.byref(tracking), ctypes.byref(r)))

    def get_nodes(self):
        """Node positions now, (x, y) of (ni+1)*(nj+1) each."""
        x = np.empty(self.grid.nnodes)
        y = np.empty(self.grid.nnodes)
        _lib.check(_lib.lib().gel_get_nodes(self._ctx, _lib.dptr(x), _lib.dptr(y)))
        return x, y

    def get_source_table(self) -> np.ndarray:
        """The source table [5][ncells] (mass, momentum x, momentum y, energy, tracer)."""
        q = np.empty((5, self.ncells))
        _lib.check(_lib.lib().gel_get_source_table(self._ctx, _lib.dptr(q)))
        return q

    def get_source_flags(self) -> np.ndarray:
        """The analytic-fallback flags [ncells] the last refresh's judge left (1: the cell
        takes the analytic closure in GEL_SRC_HYBRID)."""
        f = np.empty(self.ncells, np.int32)
        _lib.check(_lib.lib().gel_get_source_flags(self._ctx, f.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return f

    def set_source_table(self, q: np.ndarray) -> None:
        """Per-cell sources for GEL_SRC_TABLE: (mass, mom x, mom y, energy, tracer) x ncells."""
        a = np.ascontiguousarray(q, np.float64).reshape(5, self.ncells)
        _lib.check(_lib.lib().gel_set_source_table(self._ctx, _lib.dptr(a)))

    def set_rows_per_wave(self, rows: int) -> None:
        _lib.check(_lib.lib().gel_set_rows_per_wave(self._ctx, int(rows)))

    # -- sweep ------------------------------------------------------------------
    def Go_5_komponent__MK2(self, step: int) -> _lib.GelStats:
        """Advance ``step`` steps on the static grid (Setka.cpp:10121-10482)."""
        _lib.check(_lib.lib().gel_step(self._ctx, int(step), ctypes.byref(self.stats)))
        return self.stats

    def step_moving(self, x2: np.ndarray, y2: np.ndarray) -> _lib.GelStats:
        """One step with node positions after the step (Point::x2, y2, as
        Move_Setka_Calculate_2 leaves them); commits x <- x2 (Setka.cpp:10473-10480)."""
        a = np.ascontiguousarray(x2, np.float64)
        b = np.ascontiguousarray(y2, np.float64)
        _lib.check(_lib.lib().gel_step_moving(self._ctx, _lib.dptr(a), _lib.dptr(b), ctypes.byref(self.stats)))
        return self.stats

    def set_dt(self, dt: float) -> None:
        _lib.check(_lib.lib().gel_set_dt(self._ctx, float(dt)))

    def sync(self) -> _lib.GelStats:
        _lib.check(_lib.lib().gel_sync(self._ctx, ctypes.byref(self.stats)))
        return self.stats

    def enable_timing(self, on=True) -> None:
        """Event timing: False / 0 off, True / 1 around every launch, 2 around every
        Go_5_komponent__MK2 call (per-step time of back-to-back launches)."""
        _lib.check(_lib.lib().gel_enable_timing(self._ctx, int(on)))

    def kernel_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        _lib.check(_lib.lib().gel_kernel_timing(self._ctx, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


def HLLD_flux(L, R, W, n1, n2, gamma=5.0 / 3.0):
    """Batched device MHD face flux (HLLD after Solvers::HLLDQ_Korolkov metod 3,
    Solvers.cpp:1642-2185, any in-plane normal, face speed W).  L, R: [n, 9]
    primitives (ro, p, u, v, w, bx, by, bz, Q).  Returns P [n, 9] (mass, momentum
    x/y/phi, energy, B x/y/phi, tracer) and max|S| [n]."""
    L = np.ascontiguousarray(L, np.float64).reshape(-1, 9)
    R = np.ascontiguousarray(R, np.float64).reshape(-1, 9)
    n = L.shape[0]
    vec = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,))) for a in (W, n1, n2)]
    P = np.empty((n, 9))
    smax = np.empty(n)
    _lib.check(_lib.lib().gel_flux_hlld_batch(n, _lib.dptr(L), _lib.dptr(R), *[_lib.dptr(v) for v in vec],
                                              float(gamma), _lib.dptr(P), _lib.dptr(smax)))
    return P, smax


def HLLC_2d_Korolkov_b_s(L, R, W, n1, n2, rad, metod, nul_potok=False, cfl=0.1, gamma=5.0 / 3.0):
    """Batched device evaluation of the face flux (Setka.cpp:10484-10700).

    L, R: [n, 5] as (ro, Q, p, u, v) -- the argument order of the reference.
    Returns P [n, 4], PQ, Vl, Vc, Vp, dt [n] each."""
    L = np.ascontiguousarray(L, np.float64).reshape(-1, 5)
    R = np.ascontiguousarray(R, np.float64).reshape(-1, 5)
    n = L.shape[0]
    vec = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,))) for a in (W, n1, n2, rad)]
    P = np.empty((n, 4))
    outs = [np.empty(n) for _ in range(5)]
    _lib.check(_lib.lib().gel_flux_hllc2d_batch(n, _lib.dptr(L), _lib.dptr(R), *[_lib.dptr(v) for v in vec],
                                                float(cfl), float(gamma), int(metod), int(bool(nul_potok)),
                                                _lib.dptr(P), *[_lib.dptr(o) for o in outs]))
    return (P, *outs)
