"""Polar quad grids and the per-cell face tables the reference builds for them.

The reference builds its heliosphere grid from rails (Setka.cpp:19-1076) and the
per-face stencil with Setka::TVD_prepare (Setka.cpp:1294-1492).  The workloads here
are logically structured polar grids: ni cells along each ray (index i, radius) by
nj cells across the rays (index j, polar angle from the +x axis).  Node (i, j) is at
index ``i*(nj+1)+j``; cell (i, j) at ``i*nj+j`` with contour (i,j), (i+1,j),
(i+1,j+1), (i,j+1), counter-clockwise in the (x, y) half plane.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .params import (GEL_C_2, GEL_F_AXIS, GEL_F_EXTERN, GEL_F_INNER_SPHERE, GEL_F_INPUT,
                     GEL_F_USUALY)

MAXF = 8


@dataclass
class PolarGrid:
    ni: int
    nj: int
    x: np.ndarray            # (ni+1)*(nj+1) node axis coordinates
    y: np.ndarray            # (ni+1)*(nj+1) node distances from the axis
    cell_type: np.ndarray = field(default=None)   # ni*nj int8
    outer_type: np.ndarray = field(default=None)  # nj int8

    def __post_init__(self):
        self.x = np.ascontiguousarray(self.x, dtype=np.float64)
        self.y = np.ascontiguousarray(self.y, dtype=np.float64)
        assert self.x.shape == ((self.ni + 1) * (self.nj + 1),)
        if self.cell_type is None:
            self.cell_type = np.full(self.ni * self.nj, GEL_C_2, dtype=np.int8)
        self.cell_type = np.ascontiguousarray(self.cell_type, dtype=np.int8)
        if self.outer_type is None:
            self.outer_type = default_outer_type(self)
        self.outer_type = np.ascontiguousarray(self.outer_type, dtype=np.int8)

    @property
    def ncells(self) -> int:
        return self.ni * self.nj

    @property
    def nnodes(self) -> int:
        return (self.ni + 1) * (self.nj + 1)

    def node(self, i, j):
        return i * (self.nj + 1) + j

    def cell_centers(self):
        """Polygon barycentres, the formula of Cell::Get_Center (Cell.cpp:66-88)."""
        q = contour(self.ni, self.nj)
        px, py = self.x[q], self.y[q]
        A = np.zeros(self.ncells)
        gx = np.zeros(self.ncells)
        gy = np.zeros(self.ncells)
        for k in range(4):
            a, b = k, (k + 1) % 4
            cr = px[:, a] * py[:, b] - py[:, a] * px[:, b]
            A += cr
            gx += (px[:, a] + px[:, b]) * cr
            gy += (py[:, a] + py[:, b]) * cr
        A *= 0.5
        return gx / (6.0 * A), gy / (6.0 * A)


def default_outer_type(g: PolarGrid) -> np.ndarray:
    """Input where the outer face centre lies upstream (x > 0), Extern elsewhere."""
    i = g.ni
    t = np.empty(g.nj, dtype=np.int8)
    for j in range(g.nj):
        a, b = g.node(i, j), g.node(i, j + 1)
        t[j] = GEL_F_INPUT if 0.5 * (g.x[a] + g.x[b]) > 0.0 else GEL_F_EXTERN
    return t


def polar_grid(ni: int, nj: int, r_in: float, r_out: float, spacing: str = "log",
               theta0: float = 0.0, theta1: float = np.pi) -> PolarGrid:
    """Nodes at r_i (log- or uniformly spaced in [r_in, r_out]) and theta_j uniform in
    [theta0, theta1]; x = r cos(theta), y = r sin(theta).  Nodes on the axis
    (theta = 0 or pi) get y = 0 exactly."""
    i = np.arange(ni + 1, dtype=np.float64)
    if spacing == "log":
        r = r_in * (r_out / r_in) ** (i / ni)
    else:
        r = r_in + (r_out - r_in) * i / ni
    th = theta0 + (theta1 - theta0) * np.arange(nj + 1, dtype=np.float64) / nj
    R, T = np.meshgrid(r, th, indexing="ij")
    x = R * np.cos(T)
    y = R * np.sin(T)
    if theta0 == 0.0:
        y[:, 0] = 0.0
        x[:, 0] = r
    if theta1 == np.pi:
        y[:, -1] = 0.0
        x[:, -1] = -r
    return PolarGrid(ni, nj, x.ravel(), y.ravel())


def contour(ni: int, nj: int) -> np.ndarray:
    """[ncells, 4] node indices of each cell, counter-clockwise."""
    I, J = np.meshgrid(np.arange(ni), np.arange(nj), indexing="ij")
    I, J = I.ravel(), J.ravel()
    n = lambda a, b: a * (nj + 1) + b
    return np.stack([n(I, J), n(I + 1, J), n(I + 1, J + 1), n(I, J + 1)], axis=1).astype(np.int32)


@dataclass
class GeneralMesh:
    """Per-cell face slots in Cell::Grans order (slot = cell*MAXF + k): nodes A, B as
    seen from the cell, Gran_type, Sosed, Sosed_down, Sosed_up (-1 when absent).
    contour: [ncells, 4] corner nodes of quads, or with coff the corners of polygons,
    cell c's at contour[coff[c]:coff[c+1]] (Cell::contour order)."""
    contour: np.ndarray
    cell_type: np.ndarray
    nface: np.ndarray
    fa: np.ndarray
    fb: np.ndarray
    ftype: np.ndarray
    fnbr: np.ndarray
    fdown: np.ndarray
    fup: np.ndarray
    coff: np.ndarray = None


def pack_general(g: PolarGrid) -> GeneralMesh:
    """Face tables of a polar grid in the form Setka::TVD_prepare leaves them
    (Setka.cpp:1304-1434): slot k of a cell is edge contour[k] -> contour[k+1];
    Sosed_down is the cell on the far side of the master (the neighbour across the
    master's most aligned face), Sosed_up the one beyond the sosed."""
    ni, nj = g.ni, g.nj
    nc = ni * nj
    q = contour(ni, nj)
    fa = np.full(nc * MAXF, -1, np.int32)
    fb = np.full(nc * MAXF, -1, np.int32)
    ft = np.full(nc * MAXF, -1, np.int32)
    fn = np.full(nc * MAXF, -1, np.int32)
    fd = np.full(nc * MAXF, -1, np.int32)
    fu = np.full(nc * MAXF, -1, np.int32)
    cid = lambda i, j: i * nj + j if (0 <= i < ni and 0 <= j < nj) else -1
    for i in range(ni):
        for j in range(nj):
            c = i * nj + j
            # (di, dj) step across edge k, and the type when it leaves the grid
            steps = [(0, -1), (1, 0), (0, 1), (-1, 0)]
            for k, (di, dj) in enumerate(steps):
                s = c * MAXF + k
                fa[s], fb[s] = q[c, k], q[c, (k + 1) % 4]
                nb = cid(i + di, j + dj)
                fd[s] = cid(i - di, j - dj)
                if nb >= 0:
                    ft[s] = GEL_F_USUALY
                    fn[s] = nb
                    fu[s] = cid(i + 2 * di, j + 2 * dj)
                else:
                    if dj != 0:
                        ft[s] = GEL_F_AXIS
                    elif di < 0:
                        ft[s] = GEL_F_INNER_SPHERE
                    else:
                        ft[s] = int(g.outer_type[j])
    return GeneralMesh(q, g.cell_type.astype(np.int32), np.full(nc, 4, np.int32), fa, fb, ft, fn, fd, fu)


def rail_radii(t, blocks, s, R2, R3, R4):
    """Node radii along one rail at angle s with key radii R2 (inner shock), R3
    (contact), R4 (outer shock), by the spacing of Setka::Move_Setka_Calculate_2
    (Setka.cpp:4349-4380); blocks = (M1, M2, M3, M4) nodes per block as on the
    reference's rails (Rail.cpp:19-75)."""
    M1, M2, M3, M4 = blocks
    r = np.empty(M1 + M2 + M3 + M4)
    n = t.inner_rail_nodes
    for j in range(n):
        r[j] = t.r_inner + (j * 1.0 / n) ** t.p_inner * (t.r_mid - t.r_inner)
    for j in range(n, M1 - 1):
        r[j] = t.r_mid + ((j - n) * 1.0 / (M1 - 1 - n)) ** t.p_mid * (R2 - t.r_mid)
    r[M1 - 1] = R2
    for j in range(M2 - 1):
        r[M1 + j] = R2 + (R3 - R2) * (j + 1) / M2
    r[M1 + M2 - 1] = R3
    for j in range(M3 - 1):
        r[M1 + M2 + j] = R3 + (R4 - R3) * (j + 1) / M3
    r[M1 + M2 + M3 - 1] = R4
    dd = t.s_outer / np.sqrt(1.0 - 1.0 / t.p_outer)
    for j in range(M4):
        r[M1 + M2 + M3 + j] = R4 + ((j + 1) * 1.0 / M4) ** (t.p_outer * (1.0 - (s / dd) ** 2)) * (t.r_outer - R4)
    return r


def tracked_polar_grid(nj: int, blocks, key_radii, t, theta1: float = 2.0 * np.pi / 3.0):
    """A polar grid whose node columns are rails at angles theta1*c/nj and whose node
    rows M1-1, M1+M2-1, M1+M2+M3-1 are the inner shock, the contact and the outer
    shock (the reference's A rails, Rail.cpp:19-75).  key_radii(s) -> (R2, R3, R4).
    The outer-region exponent p_outer*(1 - (s/dd)^2) stays positive for
    s < dd = s_outer/sqrt(1 - 1/p_outer) (about 137 degrees with the reference's
    constants), which bounds theta1.  Returns (grid, rail angles, rows)."""
    M1, M2, M3, M4 = blocks
    ni = M1 + M2 + M3 + M4 - 1
    rails = theta1 * np.arange(nj + 1, dtype=np.float64) / nj
    dd = t.s_outer / np.sqrt(1.0 - 1.0 / t.p_outer)
    if not theta1 < dd:
        raise ValueError(f"rails beyond {dd:.4f} rad have a non-positive outer exponent")
    x = np.empty((ni + 1, nj + 1))
    y = np.empty((ni + 1, nj + 1))
    for c, s in enumerate(rails):
        R2, R3, R4 = key_radii(s)
        r = rail_radii(t, blocks, s, R2, R3, R4)
        x[:, c] = r * np.cos(s)
        y[:, c] = r * np.sin(s)
    y[:, 0] = 0.0
    g = PolarGrid(ni, nj, x.ravel(), y.ravel())
    rows = (M1 - 1, M1 + M2 - 1, M1 + M2 + M3 - 1)
    return g, rails, rows
