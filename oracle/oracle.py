"""ctypes front end of the CPU oracle (oracle/libgelios_oracle.so).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the CPU-baseline
leg of bench.py.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from gelios_2d_amd.grid import MAXF, GeneralMesh, PolarGrid, pack_general
from gelios_2d_amd.params import GEL_MC_NFIELDS, GelMcParams, GelMcStats, GelParams, GelTracking

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgelios_oracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OraMesh(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int), ("n_cells", ctypes.c_int),
                ("x", _dp), ("y", _dp), ("x2", _dp), ("y2", _dp),
                ("contour", _ip), ("cell_type", _ip), ("nface", _ip),
                ("fa", _ip), ("fb", _ip), ("ftype", _ip), ("fnbr", _ip), ("fdown", _ip), ("fup", _ip),
                ("ptype", _ip), ("coff", _ip)]


class OraMhd(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("ro", "p", "u", "v", "w", "bx", "by", "bz", "Q")]


class OraGas(ctypes.Structure):
    _fields_ = [("ro", ctypes.c_double), ("p", ctypes.c_double), ("u", ctypes.c_double),
                ("v", ctypes.c_double), ("Q", ctypes.c_double)]


def build(quiet: bool = True) -> None:
    """Compile the restatement (and the reference harness where the reference
    sources exist) with oracle/Makefile."""
    subprocess.run(["make", "-C", HERE, "-j8"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ora_hllc2d.restype = ctypes.c_double
        L.ora_hllc2d.argtypes = [ctypes.POINTER(OraGas), ctypes.POINTER(OraGas), ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        for name, n in (("ora_minmod", 2), ("ora_linear3", 7), ("ora_linear2", 5), ("ora_polar_angle", 2)):
            f = getattr(L, name)
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.c_double] * n
        L.ora_polar_perenos.restype = None
        L.ora_polar_perenos.argtypes = [ctypes.c_double] * 4 + [_dp, _dp]
        L.ora_cell_center.argtypes = [_dp, _dp, _ip, _dp, _dp]
        L.ora_cell_area.restype = ctypes.c_double
        L.ora_cell_area.argtypes = [_dp, _dp, _ip]
        L.ora_cell_volume_rotate.restype = ctypes.c_double
        L.ora_cell_volume_rotate.argtypes = [_dp, _dp, _ip, ctypes.c_double]
        L.ora_source_mk1.argtypes = [ctypes.POINTER(GelParams), _dp, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        for name in ("ora_step_mk2", "ora_step_mhd"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.POINTER(OraMesh), ctypes.POINTER(GelParams), _dp, _dp, _dp,
                          ctypes.c_double, _dp, ctypes.POINTER(ctypes.c_longlong),
                          ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
        L.ora_godunov.restype = ctypes.c_double
        L.ora_godunov.argtypes = [_dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
        L.ora_step_inner.restype = ctypes.c_int
        L.ora_step_inner.argtypes = [ctypes.POINTER(OraMesh), ctypes.POINTER(GelParams), _dp, _dp, _dp,
                                     ctypes.c_double, _dp, ctypes.POINTER(ctypes.c_longlong),
                                     ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
        L.ora_step_mf.restype = ctypes.c_int
        L.ora_step_mf.argtypes = [ctypes.POINTER(OraMesh), ctypes.POINTER(GelParams), _dp, _dp, ctypes.c_double, _dp,
                                  ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
        L.ora_track_nodes.restype = ctypes.c_int
        L.ora_track_nodes.argtypes = [ctypes.POINTER(GelTracking), ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                      ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.ora_mhd_face.restype = ctypes.c_int
        L.ora_mhd_face.argtypes = [ctypes.POINTER(OraMesh), ctypes.POINTER(GelParams), _dp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, _dp, _dp, _dp]
        L.ora_set_face_once.restype = None
        L.ora_set_face_once.argtypes = [ctypes.c_int]
        L.ora_hlld.restype = ctypes.c_double
        L.ora_hlld.argtypes = [ctypes.POINTER(OraMhd), ctypes.POINTER(OraMhd), ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, _dp]
        L.ora_sensor_draw.restype = ctypes.c_double
        L.ora_sensor_draw.argtypes = [_ip]
        L.ora_mc_seed_primary.restype = None
        L.ora_mc_seed_primary.argtypes = [ctypes.c_ulonglong, ctypes.c_longlong, _ip]
        L.ora_mc_velosity.restype = None
        L.ora_mc_velosity.argtypes = [ctypes.c_double, ctypes.c_double, _dp]
        L.ora_mc_int.restype = None
        L.ora_mc_int.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.ora_mc_exchange.restype = ctypes.c_int
        L.ora_mc_exchange.argtypes = [_ip, _dp, ctypes.c_double, _dp]
        L.ora_mc_inflow.restype = ctypes.c_double
        L.ora_mc_inflow.argtypes = [ctypes.c_double]
        L.ora_mc_run.restype = ctypes.c_int
        L.ora_mc_run.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int8),
                                 ctypes.POINTER(GelMcParams), _dp, ctypes.POINTER(GelMcStats)]
        L.ora_mc_run_mesh.restype = ctypes.c_int
        L.ora_mc_run_mesh.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _ip, _ip, _ip, _ip, _ip, _ip, _dp,
                                      ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(GelMcParams), _dp,
                                      ctypes.POINTER(GelMcStats)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def hllc2d(L, R, W, n1, n2, rad, cfl, gamma, metod, nul_potok):
    """Point-wise flux. L, R = (ro, Q, p, u, v) (argument order of
    HLLC_2d_Korolkov_b_s). Returns (P[4], PQ, Vl, Vc, Vp, dt)."""
    gl = OraGas(L[0], L[2], L[3], L[4], L[1])
    gr = OraGas(R[0], R[2], R[3], R[4], R[1])
    P = np.zeros(4)
    o = [ctypes.c_double() for _ in range(4)]
    t = lib().ora_hllc2d(ctypes.byref(gl), ctypes.byref(gr), W, n1, n2, rad, cfl, gamma, int(metod),
                         int(nul_potok), _d(P), *[ctypes.byref(v) for v in o])
    return P, o[0].value, o[1].value, o[2].value, o[3].value, t


def hlld(L, R, W, n1, n2, gamma=5.0 / 3.0):
    """Point-wise HLLD. L, R = 9 primitives (ro, p, u, v, w, bx, by, bz, Q).
    Returns (P[9] = mass, momentum x/y/phi, energy, B x/y/phi, tracer; max|S|)."""
    P = np.zeros(9)
    a, b = OraMhd(*map(float, L)), OraMhd(*map(float, R))
    uu = lib().ora_hlld(ctypes.byref(a), ctypes.byref(b), W, n1, n2, gamma, _d(P))
    return P, uu


def godunov(q1, q2, n, w=0.0, dist=1.0, cfl=0.1, gamma=5.0 / 3.0):
    """Point-wise exact Riemann solver.  q1, q2 = (ro, p, u, v, w); n = (n1, n2, n3).
    Returns (q[5] fluxes, (dsl, dsp, dsc), time, status)."""
    a = np.ascontiguousarray(q1, np.float64)
    b = np.ascontiguousarray(q2, np.float64)
    nn = np.ascontiguousarray(n, np.float64)
    q = np.zeros(5)
    sl, sp, sc = np.zeros(1), np.zeros(1), np.zeros(1)
    st = ctypes.c_int(0)
    t = lib().ora_godunov(_d(a), _d(b), _d(nn), float(w), float(dist), float(cfl), float(gamma), _d(q), _d(sl),
                          _d(sp), _d(sc), ctypes.byref(st))
    return q, (sl[0], sp[0], sc[0]), t, st.value


class OracleSweep:
    """Holds a general mesh and steps the restated Go_5_komponent__MK2 cell loop."""

    def __init__(self, mesh: GeneralMesh, x, y, prm: GelParams, x2=None, y2=None):
        self.mesh = mesh
        self.prm = prm.copy()
        self.x = np.ascontiguousarray(x, np.float64)
        self.y = np.ascontiguousarray(y, np.float64)
        self.x2 = np.ascontiguousarray(self.x if x2 is None else x2, np.float64)
        self.y2 = np.ascontiguousarray(self.y if y2 is None else y2, np.float64)
        self.nc = len(mesh.cell_type)
        self.nv = {1: 9, 2: 21}.get(prm.physics, 5)
        self._keep = [np.ascontiguousarray(a, np.int32) for a in
                      (mesh.contour.ravel(), mesh.cell_type, mesh.nface, mesh.fa, mesh.fb, mesh.ftype,
                       mesh.fnbr, mesh.fdown, mesh.fup)]
        self._coff = None if mesh.coff is None else np.ascontiguousarray(mesh.coff, np.int32)
        self.n_rho_clamp = ctypes.c_longlong(0)
        self.n_p_clamp = ctypes.c_longlong(0)
        self.dt = prm.dt0
        self.src = None
        self.ptype = None

    @classmethod
    def from_polar(cls, g: PolarGrid, prm: GelParams):
        return cls(pack_general(g), g.x, g.y, prm)

    def _mesh(self):
        k = self._keep
        pt = _i(self.ptype) if self.ptype is not None else None
        co = _i(self._coff) if self._coff is not None else None
        return OraMesh(len(self.x), self.nc, _d(self.x), _d(self.y), _d(self.x2), _d(self.y2),
                       *[_i(a) for a in k], pt, co)

    def set_nodes(self, x, y, x2, y2):
        self.x = np.ascontiguousarray(x, np.float64)
        self.y = np.ascontiguousarray(y, np.float64)
        self.x2 = np.ascontiguousarray(x2, np.float64)
        self.y2 = np.ascontiguousarray(y2, np.float64)

    def step(self, prim, nthreads=0, dt_prev=None):
        """One step from prim (5 x nc); returns (next, status).  Uses and updates the
        lagged time step unless dt_prev is given."""
        now = np.ascontiguousarray(prim, np.float64).reshape(self.nv, self.nc)
        nxt = now.copy()
        dt_new = ctypes.c_double()
        m = self._mesh()
        src = _d(self.src) if self.src is not None else None
        dtp = self.dt if dt_prev is None else dt_prev
        if self.nv == 21:
            st = lib().ora_step_mf(ctypes.byref(m), ctypes.byref(self.prm), _d(now), _d(nxt), dtp, ctypes.byref(dt_new),
                                   ctypes.byref(self.n_rho_clamp), ctypes.byref(self.n_p_clamp), nthreads)
        else:
            fn = lib().ora_step_mhd if self.nv == 9 else lib().ora_step_mk2
            if self.prm.physics == 3:
                fn = lib().ora_step_inner
            st = fn(ctypes.byref(m), ctypes.byref(self.prm), _d(now), _d(nxt), src, dtp, ctypes.byref(dt_new),
                    ctypes.byref(self.n_rho_clamp), ctypes.byref(self.n_p_clamp), nthreads)
        if dt_prev is None:
            self.dt = dt_new.value
        self.last_dt_new = dt_new.value
        return nxt, st

    def mhd_face(self, prim, c, k, dt_prev=None):
        """Face slot k of cell c as the MHD cell loop evaluates it from that cell:
        (flux x area products[9], largest wave speed, (L[9], R[9], W, n1, n2, S))."""
        now = np.ascontiguousarray(prim, np.float64).reshape(self.nv, self.nc)
        PS = np.zeros(9)
        lrw = np.zeros(22)
        uu = ctypes.c_double()
        m = self._mesh()
        st = lib().ora_mhd_face(ctypes.byref(m), ctypes.byref(self.prm), _d(now), int(c), int(k),
                                self.dt if dt_prev is None else dt_prev, _d(PS), ctypes.byref(uu), _d(lrw))
        if st not in (0, 6):
            raise RuntimeError(f"ora_mhd_face failed with status {st}")
        return PS, uu.value, (lrw[:9], lrw[9:18], lrw[18], lrw[19], lrw[20], lrw[21])

    def run_motion(self, prim, steps, x0, y0, amp, period, phase, nthreads=0):
        """Steps on the prescribed breathing grid: before step n the nodes are at
        base * s[n mod period], after it at base * s[(n+1) mod period]."""
        from gelios_2d_amd.configs import motion_scales
        s = motion_scales(amp, period, phase)
        x0 = np.asarray(x0, np.float64)
        y0 = np.asarray(y0, np.float64)
        cur = np.ascontiguousarray(prim, np.float64).reshape(self.nv, self.nc)
        for n in range(steps):
            so, sn = s[n % period], s[(n + 1) % period]
            self.set_nodes(x0 * so, y0 * so, x0 * sn, y0 * sn)
            cur, st = self.step(cur, nthreads)
            if st != 0:
                raise RuntimeError(f"oracle step failed with status {st}")
        return cur

    def run(self, prim, steps, nthreads=0):
        cur = np.ascontiguousarray(prim, np.float64).reshape(self.nv, self.nc)
        for _ in range(steps):
            cur, st = self.step(cur, nthreads)
            if st != 0:
                raise RuntimeError(f"oracle step failed with status {st}")
        return cur


def track_nodes(t: GelTracking, ni, nj, x, y, prim, rails, gamma, dt):
    """Node positions after a step of the fitted-surface tracking (Move_surface and the
    A rails of Move_Setka_Calculate_2).  Returns (x2, y2, status)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    now = np.ascontiguousarray(prim, np.float64)
    rails = np.ascontiguousarray(rails, np.float64)
    x2, y2 = np.empty_like(x), np.empty_like(y)
    st = lib().ora_track_nodes(ctypes.byref(t), ni, nj, _d(x), _d(y), _d(now), _d(rails), float(gamma), float(dt),
                               _d(x2), _d(y2))
    return x2, y2, st


def surface_ptype(g: PolarGrid, rows):
    """Node types of the fitted surfaces: 2 on the inner-shock row, 1 on the contact
    row, 3 on the outer-shock row (P_Inner_shock, P_Contact, P_Outer_shock)."""
    pt = np.zeros(g.nnodes, np.int32)
    for code, r in zip((2, 1, 3), rows):
        pt[r * (g.nj + 1): (r + 1) * (g.nj + 1)] = code
    return pt


class TrackedOracle:
    """The restated Go_5_komponent__MK2 with fitted-surface tracking on a polar grid:
    every step moves the surface nodes and redistributes the rails, sweeps with the
    exact solver on the surface faces, then commits the nodes."""

    def __init__(self, g: PolarGrid, prm: GelParams, t: GelTracking, rails):
        self.g = g
        self.t = t.copy()
        self.rails = np.ascontiguousarray(rails, np.float64)
        self.sw = OracleSweep(pack_general(g), g.x, g.y, prm)
        self.sw.ptype = surface_ptype(g, t.rows)
        self.x = g.x.copy()
        self.y = g.y.copy()

    def run(self, prim, steps, nthreads=0):
        cur = np.ascontiguousarray(prim, np.float64).reshape(5, self.g.ncells)
        for _ in range(steps):
            x2, y2, st = track_nodes(self.t, self.g.ni, self.g.nj, self.x, self.y, cur, self.rails,
                                     self.sw.prm.gamma, self.sw.dt)
            if st != 0:
                raise RuntimeError(f"oracle tracking failed with status {st}")
            self.sw.set_nodes(self.x, self.y, x2, y2)
            cur, st = self.sw.step(cur, nthreads)
            if st != 0:
                raise RuntimeError(f"oracle step failed with status {st}")
            self.x, self.y = x2, y2
        return cur


# ---------------------------------------------------------------- Monte Carlo ----
def sensor_draws(seed, n):
    """n draws of the reference's random sensor (sensor.cpp:11-31) from seed (a1, a2, a3)."""
    s = np.ascontiguousarray(seed, np.int32).copy()
    return np.array([lib().ora_sensor_draw(_i(s)) for _ in range(n)])


def mc_rates(u, cp, sig_a=0.1307345665):
    """[n][6]: Velosity_1..3 and int_1..3 (int_k 0 where u / cp > 7)."""
    u = np.atleast_1d(np.asarray(u, np.float64))
    cp = np.broadcast_to(np.asarray(cp, np.float64), u.shape)
    out = np.zeros((len(u), 6))
    o = np.zeros(3)
    for k in range(len(u)):
        lib().ora_mc_velosity(float(u[k]), float(cp[k]), _d(o))
        out[k, :3] = o
        if u[k] / cp[k] <= 7.0:
            lib().ora_mc_int(float(u[k]), float(cp[k]), float(sig_a), _d(o))
            out[k, 3:] = o
    return out


def mc_exchange(inp, seeds, reps, sig_a=0.1307345665):
    """Change_Velosity4 with I = 0 from seeded sensors: inp[n][7], seeds[n][3] ->
    (out[n][reps][4] = Wr Wthe Wphi mu, state[n][3])."""
    inp = np.ascontiguousarray(inp, np.float64)
    n = len(inp)
    out = np.zeros((n, reps, 4))
    state = np.zeros((n, 3), np.int32)
    for k in range(n):
        s = np.ascontiguousarray(seeds[k], np.int32).copy()
        row = np.ascontiguousarray(inp[k])
        for r in range(reps):
            o = np.zeros(4)
            lib().ora_mc_exchange(_i(s), _d(row), float(sig_a), _d(o))
            out[k, r] = o
        state[k] = s
    return out, state


def mc_run(g: PolarGrid, prim, pop, prm: GelMcParams):
    """The tracer on a polar grid; returns (out[GEL_MC_NFIELDS][ncells], stats dict)."""
    nc = g.ni * g.nj
    x = np.ascontiguousarray(g.x, np.float64)
    y = np.ascontiguousarray(g.y, np.float64)
    prim = np.ascontiguousarray(prim, np.float64)
    pop = np.ascontiguousarray(pop, np.int8)
    out = np.zeros((GEL_MC_NFIELDS, nc))
    st = GelMcStats()
    lib().ora_mc_run(g.ni, g.nj, _d(x), _d(y), _d(prim), pop.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                     ctypes.byref(prm), _d(out), ctypes.byref(st))
    return out, st.as_dict()


def mc_mesh_slots(mesh):
    """The face slots of a general mesh (gelios_2d_amd.mesh.Mesh) as the tracer walks them,
    compacted in Cell::Grans order: (soff, sA, sB, scode).  scode is the way on: Sosed
    where the slot has one (any type: the block grid's Inner_sphere faces lead to the
    central cells), -3 for Axis (the reference's tracer skips them, Setka.cpp:12897),
    -2 for an Inner_sphere slot with no Sosed (crossed in a straight line), -1 for any
    other boundary (the atom leaves; primaries enter there)."""
    nc = len(mesh.nface)
    maxf = len(mesh.fa) // nc
    keep = (np.arange(maxf)[None, :] < np.asarray(mesh.nface)[:, None]).ravel()
    sA = np.ascontiguousarray(np.asarray(mesh.fa)[keep], np.int32)
    sB = np.ascontiguousarray(np.asarray(mesh.fb)[keep], np.int32)
    t = np.asarray(mesh.ftype)[keep]
    n = np.asarray(mesh.fnbr)[keep]
    code = np.where(n >= 0, n, np.where(t == 1, -2, -1))
    code = np.where(t == 4, -3, code).astype(np.int32)
    soff = np.zeros(nc + 1, np.int32)
    soff[1:] = np.cumsum(mesh.nface)
    return soff, sA, sB, np.ascontiguousarray(code)


def mc_run_mesh(mesh, prim, pop, prm: GelMcParams):
    """The tracer on a general mesh; returns (out[GEL_MC_NFIELDS][ncells], stats dict)."""
    nc = len(mesh.nface)
    soff, sA, sB, code = mc_mesh_slots(mesh)
    x = np.ascontiguousarray(mesh.x, np.float64)
    y = np.ascontiguousarray(mesh.y, np.float64)
    coff = np.ascontiguousarray(mesh.coff, np.int32)
    corners = np.ascontiguousarray(mesh.corners, np.int32)
    prim = np.ascontiguousarray(prim, np.float64)
    pop = np.ascontiguousarray(pop, np.int8)
    out = np.zeros((GEL_MC_NFIELDS, nc))
    st = GelMcStats()
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))   # noqa: E731
    lib().ora_mc_run_mesh(nc, len(x), _d(x), _d(y), ip(coff), ip(corners), ip(soff), ip(sA), ip(sB), ip(code),
                          _d(prim), pop.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), ctypes.byref(prm), _d(out),
                          ctypes.byref(st))
    return out, st.as_dict()


def mc_population(cell_type, nc):
    """Population of a secondary born in a cell (Setka.cpp:13384-13398): C_5 -> 3, C_4 -> 2,
    C_1 and C_centr -> 0, others 1."""
    if cell_type is None:
        return np.ones(nc, np.int8)
    ct = np.asarray(cell_type)
    pop = np.ones(nc, np.int8)
    pop[ct == 4] = 3
    pop[ct == 3] = 2
    pop[(ct == 0) | (ct == 5)] = 0
    return pop
