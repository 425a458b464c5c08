"""The reference's own grid (the block-structured heliosphere grid of Setka, with its
four rail families and central wedge cells, Setka.cpp:19-1076) as face tables for the
general-mesh sweep of libgelios.so.

A :class:`Mesh` holds what the reference's cell loop reads (Setka.cpp:10148-10470):
node positions and fitted-surface point types, cell contours and types, and for every
cell its faces in Cell::Grans order with the nodes A, B as seen from the cell, the face
type, Sosed, and the stencil cells Sosed_down / Sosed_up that Setka::TVD_prepare
(Setka.cpp:1294-1434) leaves.  :func:`mesh_from_alpha` builds it from a checkpoint
(Setka::Save_Setka_ALL_ALPHA format, as :func:`checkpoint.read_alpha` parses it) the
way Setka::Download_Setka_ALL_ALPHA_2_0 (Setka.cpp:2899-3258) and TVD_prepare do.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .grid import MAXF, GeneralMesh
from .params import GEL_F_USUALY

# Point_type (Point.h:7-21) -> fitted-surface codes of the sweep
_P_CONTACT, _P_INNER_SHOCK, _P_OUTER_SHOCK = 6, 7, 8


@dataclass
class Mesh:
    x: np.ndarray             # [nn] Point::x
    y: np.ndarray             # [nn] Point::y
    ptype: np.ndarray         # [nn] int32: 0, 1 P_Contact, 2 P_Inner_shock, 3 P_Outer_shock
    coff: np.ndarray          # [nc + 1] corner offsets
    corners: np.ndarray       # Cell::contour of every cell, concatenated
    cell_type: np.ndarray     # [nc] Cell_type
    nface: np.ndarray         # [nc]
    fa: np.ndarray            # [nc * MAXF] face slots in Cell::Grans order
    fb: np.ndarray
    ftype: np.ndarray
    fnbr: np.ndarray
    fdown: np.ndarray
    fup: np.ndarray

    @property
    def nnodes(self) -> int:
        return len(self.x)

    @property
    def ncells(self) -> int:
        return len(self.cell_type)

    def general(self) -> GeneralMesh:
        """The same tables in the oracle's form."""
        return GeneralMesh(self.corners, self.cell_type.astype(np.int32), self.nface, self.fa, self.fb, self.ftype,
                           self.fnbr, self.fdown, self.fup, coff=self.coff)

    def cell_centers(self, x=None, y=None):
        """Cell::Get_Center (Cell.cpp:66-88) of every cell at node positions x, y."""
        x = self.x if x is None else x
        y = self.y if y is None else y
        return _centers(self.coff, self.corners, x, y)


def _centers(coff, corners, x, y):
    nc = len(coff) - 1
    cx = np.empty(nc)
    cy = np.empty(nc)
    for c in range(nc):
        q = corners[coff[c]:coff[c + 1]]
        n = len(q)
        A = Gx = Gy = 0.0
        for k in range(n):
            a, b = q[k], q[k + 1 if k + 1 < n else 0]
            cr = x[a] * y[b] - y[a] * x[b]
            A = A + cr
            Gx = Gx + (x[a] + x[b]) * cr
            Gy = Gy + (y[a] + y[b]) * cr
        A = A * 0.5
        cx[c] = Gx / (6.0 * A)
        cy[c] = Gy / (6.0 * A)
    return cx, cy


def tvd_stencils(x, y, fA, fB, ftype, fsosed, fcopy, cell_faces, ccx, ccy):
    """Sosed_up and Sosed_down of every face, as Setka::TVD_prepare computes them
    (Setka.cpp:1304-1434).  Faces are numbered as after Download_Setka_ALL_ALPHA_2_0:
    the main faces first, then the copies.  fsosed[f]: Gran::Sosed (-1 none);
    fcopy[f]: the number of Gran_copy (-1 none); cell_faces[c]: Cell::Grans as face
    numbers; ccx, ccy: cell centres.  For every face with a Sosed: the face of the Sosed
    (other than the face's own copy) whose direction from the Sosed's centre is most
    aligned with the direction from the face centre to the Sosed's centre; Sosed_up is
    the cell across it.  Sosed_down is the copy's Sosed_up."""
    nf = len(fA)
    fx = [(x[fA[f]] + x[fB[f]]) / 2.0 for f in range(nf)]      # Gran::Get_Center
    fy = [(y[fA[f]] + y[fB[f]]) / 2.0 for f in range(nf)]
    up = [-1] * nf
    for i in range(nf):
        s = fsosed[i]
        if s < 0:
            continue
        x3, y3 = ccx[s], ccy[s]
        n1, n2 = x3 - fx[i], y3 - fy[i]
        n = math.sqrt(n1 * n1 + n2 * n2)
        n1, n2 = n1 / n, n2 / n
        N, B = -1.0, -1
        for j in cell_faces[s]:
            if ftype[j] == GEL_F_USUALY and fcopy[j] == i:
                continue
            a, b = fx[j] - x3, fy[j] - y3
            n = math.sqrt(a * a + b * b)
            a, b = a / n, b / n
            if a * n1 + b * n2 > N:
                N, B = a * n1 + b * n2, j
        up[i] = fsosed[B]
    down = [up[fcopy[i]] if fcopy[i] >= 0 else -1 for i in range(nf)]
    return up, down


@dataclass
class MeshRails:
    """The reference's rails and fitted-surface lines of a block grid, by node and cell
    index: Setka::A_Rails, B_Rails, C_Rails, D_Rails (Rail.h: type, s, M1..M4,
    All_point, Key_point) and Line_Contact, Line_Inner, Line_Outer (per face Gran::A, B,
    Master, Sosed), for gel_set_mesh_tracking."""
    kind: np.ndarray          # [nr] 0 A, 1 B, 2 C, 3 D
    m: np.ndarray             # [nr, 4] M1..M4
    s: np.ndarray             # [nr]
    all_offset: np.ndarray    # [nr + 1]
    all: np.ndarray           # concatenated All_point
    key: np.ndarray           # [nr, 4] (-1 where the rail has fewer)
    lines: tuple              # three [n, 4] arrays: A, B, Master, Sosed
    left: float = -1500.0 / 390.222          # Left_ (Help.h:85)
    d_chain: int = 13                        # n_outer_shock (Help.h:68)


def mesh_from_alpha(m) -> Mesh:
    """The grid of a checkpoint (:class:`checkpoint.AlphaMesh`) as the reference holds it
    after Download_Setka_ALL_ALPHA_2_0 and TVD_prepare."""
    nmain, ncopy = len(m.gran), len(m.gran_copy)
    nc = len(m.cell_type)
    fA = np.concatenate([m.gran[:, 0], m.gran_copy[:, 0]]).astype(np.int64)
    fB = np.concatenate([m.gran[:, 1], m.gran_copy[:, 1]]).astype(np.int64)
    ftype = np.concatenate([m.gran[:, 2], m.gran_copy[:, 2]]).astype(np.int64)
    # links (Setka.cpp:3104-3143): main (master, sosed, copy), copy (master, sosed, main)
    fsosed = np.concatenate([m.gran_links[:, 1], np.where(m.copy_links[:, 1] >= 0, m.copy_links[:, 1], -1)])
    fcopy = np.concatenate([np.where(m.gran_links[:, 2] >= 0, nmain + m.gran_links[:, 2], -1),
                            np.where(m.copy_links[:, 1] >= 0, m.copy_links[:, 2], -1)])
    cell_faces = [[k if main else nmain + k for k, main in g] for g in m.cell_grans]
    coff = np.zeros(nc + 1, np.int32)
    coff[1:] = np.cumsum([len(c) for c in m.contour])
    corners = np.array([p for c in m.contour for p in c], np.int32)
    x, y = m.px.astype(np.float64), m.py.astype(np.float64)
    ccx, ccy = _centers(coff, corners, x, y)
    up, down = tvd_stencils(x, y, fA, fB, ftype, fsosed, fcopy, cell_faces, ccx, ccy)
    nface = np.array([len(g) for g in cell_faces], np.int32)
    if nface.max() > MAXF:
        raise ValueError(f"a cell has {nface.max()} faces; the face tables hold {MAXF}")
    slots = {k: np.full(nc * MAXF, -1, np.int32) for k in ("fa", "fb", "ftype", "fnbr", "fdown", "fup")}
    for c, g in enumerate(cell_faces):
        for k, f in enumerate(g):
            s = c * MAXF + k
            slots["fa"][s], slots["fb"][s], slots["ftype"][s] = fA[f], fB[f], ftype[f]
            slots["fnbr"][s], slots["fdown"][s], slots["fup"][s] = fsosed[f], down[f], up[f]
    pt = np.zeros(len(x), np.int32)
    pt[m.ptype == _P_CONTACT] = 1
    pt[m.ptype == _P_INNER_SHOCK] = 2
    pt[m.ptype == _P_OUTER_SHOCK] = 3
    return Mesh(x, y, pt, coff, corners, m.cell_type.astype(np.int32), nface, **slots)


def mesh_from_slots(x, y, ptype, ncorner, corners, cell_type, nface, slots) -> Mesh:
    """A mesh from the reference's objects flattened: node positions and point types
    (0, 1 contact, 2 inner shock, 3 outer shock), each cell's contour and type, and for
    each cell its faces in Cell::Grans order as [A, B, type, Sosed, Sosed_down,
    Sosed_up] rows (slots [nc][MAXF][6], -1 past nface)."""
    nc = len(cell_type)
    coff = np.zeros(nc + 1, np.int32)
    coff[1:] = np.cumsum(ncorner)
    sl = np.asarray(slots, np.int32).reshape(nc, MAXF, 6)
    cols = {k: np.ascontiguousarray(sl[:, :, i].ravel(), np.int32)
            for i, k in enumerate(("fa", "fb", "ftype", "fnbr", "fdown", "fup"))}
    return Mesh(np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64),
                np.ascontiguousarray(ptype, np.int32), coff, np.ascontiguousarray(corners, np.int32),
                np.ascontiguousarray(cell_type, np.int32), np.ascontiguousarray(nface, np.int32), **cols)


def mesh_from_polar(g) -> Mesh:
    """A polar grid (:class:`~gelios_2d_amd.grid.PolarGrid`) in the general form, with the
    face tables of grid.pack_general (TVD_prepare's on such grids)."""
    from .grid import pack_general
    gm = pack_general(g)
    nc = g.ncells
    return Mesh(np.ascontiguousarray(g.x, np.float64), np.ascontiguousarray(g.y, np.float64),
                np.zeros(g.nnodes, np.int32), np.arange(0, 4 * nc + 1, 4, dtype=np.int32),
                np.ascontiguousarray(gm.contour.ravel(), np.int32), g.cell_type.astype(np.int32),
                np.ascontiguousarray(gm.nface, np.int32), *[np.ascontiguousarray(getattr(gm, k), np.int32)
                                                           for k in ("fa", "fb", "ftype", "fnbr", "fdown", "fup")])


def state_from_alpha(m) -> np.ndarray:
    """The plasma state (ro, p, u, v, Q) x ncells of a checkpoint."""
    return np.ascontiguousarray(m.state[:, 0:5].T, np.float64)
