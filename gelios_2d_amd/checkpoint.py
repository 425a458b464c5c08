"""The reference's checkpoint format (Setka::Save_Setka_ALL_ALPHA, Setka.cpp:2355-2574;
read back by Setka::Download_Setka_ALL_ALPHA_2_0, Setka.cpp:2899-3258).

A checkpoint is a whitespace-separated text file holding the whole mesh (points, main
and copy faces, the four rail families, cells with their contours and faces, the face
links, the cells around every point, the three fitted-surface lines) and one state
buffer of every cell (plasma, four hydrogen fluids, the I_* moments, the Monte Carlo
moments of four populations and the k factors), numbers printed as a C++ stream prints
a double by default (six significant digits, ``%g``).

:func:`read_alpha` parses any such file; :func:`polar_from_alpha` turns one whose cells
form a logically structured polar grid into a :class:`~gelios_2d_amd.grid.PolarGrid`
and the 21-field state and 19 source inputs of this library; :func:`alpha_from_polar`
and :func:`write_alpha` write a polar grid and a state in the format, laid out as the
reference lays out that grid (faces and copies in creation order, boundary faces with
the copy's master absent).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .grid import PolarGrid
from .params import GEL_F_AXIS, GEL_F_INNER_SPHERE, GEL_F_USUALY


@dataclass
class AlphaMesh:
    header: list                      # N1 N2 N3 N4 M1 M2 M3 M4
    px: np.ndarray                    # points
    py: np.ndarray
    ptype: np.ndarray
    gran: np.ndarray                  # [n][3] main faces: A, B, type
    gran_copy: np.ndarray             # [n][3] copies: A, B, type
    rails: list                       # 4 lists (A, B, C, D) of dict(m=[M1..M4], s, type, all, key)
    cell_type: np.ndarray
    contour: list                     # per cell: point numbers
    cell_grans: list                  # per cell: [(face number, main flag)]
    gran_links: np.ndarray            # [n][3] master, sosed, copy
    copy_links: np.ndarray            # [n][3] master, sosed, copy
    point_cells: list                 # per point: cell numbers
    lines: list                       # contact, inner, outer: main face numbers
    state: np.ndarray                 # [ncells][44] in file order
    field_names: list = field(default_factory=lambda: STATE_FIELDS)


STATE_FIELDS = (["ro", "p", "u", "v", "Q"]
                + [f"{n}_H{k}" for k in range(1, 5) for n in ("ro", "p", "u", "v")]
                + ["I_u", "I_v", "I_T"]
                + [f"H_{n}[{j}]" for j in range(4) for n in ("n", "u", "v", "T")]
                + ["k_u", "k_v", "k_T"])


def _tokens(path):
    with open(path, "r") as f:
        return f.read().split()


def read_alpha(path) -> AlphaMesh:
    """Parse a checkpoint in the order Save_Setka_ALL_ALPHA writes it."""
    t = _tokens(path)
    pos = 0

    def take(n=1, conv=float):
        nonlocal pos
        v = [conv(x) for x in t[pos:pos + n]]
        if len(v) != n:
            raise ValueError(f"{path}: truncated at token {pos}")
        pos += n
        return v if n > 1 else v[0]

    header = [take(1, int) for _ in range(8)]
    n = take(1, int)
    pts = np.array([take(3, float) for _ in range(n)]).reshape(n, 3)
    n = take(1, int)
    gran = np.array([take(3, int) for _ in range(n)], dtype=np.int64).reshape(n, 3)
    n = take(1, int)
    gran_copy = np.array([take(3, int) for _ in range(n)], dtype=np.int64).reshape(n, 3)
    rails = []
    for _ in range(4):
        fam = []
        for _ in range(take(1, int)):
            m = [take(1, int) for _ in range(4)]
            s = take(1, float)
            ty = take(1, int)
            all_p = [take(1, int) for _ in range(take(1, int))]
            key = [take(1, int) for _ in range(take(1, int))]
            fam.append(dict(m=m, s=s, type=ty, all=all_p, key=key))
        rails.append(fam)
    nc = take(1, int)
    ctype, contour, cgr = [], [], []
    for _ in range(nc):
        ctype.append(take(1, int))
        contour.append([take(1, int) for _ in range(take(1, int))])
        cgr.append([(take(1, int), take(1, int)) for _ in range(take(1, int))])
    n = take(1, int)
    gl = np.array([take(3, int) for _ in range(n)], dtype=np.int64).reshape(n, 3)
    n = take(1, int)
    cl = np.array([take(3, int) for _ in range(n)], dtype=np.int64).reshape(n, 3)
    n = take(1, int)
    pc = [[take(1, int) for _ in range(take(1, int))] for _ in range(n)]
    lines = [[take(1, int) for _ in range(take(1, int))] for _ in range(3)]
    n = take(1, int)
    state = np.array([take(len(STATE_FIELDS), float) for _ in range(n)]).reshape(n, len(STATE_FIELDS))
    if pos != len(t):
        raise ValueError(f"{path}: {len(t) - pos} tokens after the cell states")
    return AlphaMesh(header, pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].astype(np.int64), gran, gran_copy,
                     rails, np.array(ctype, np.int64), contour, cgr, gl, cl, pc, lines, state)


def _g(v) -> str:
    """A double as a C++ stream prints it by default."""
    return f"{float(v):g}"


def write_alpha(path, m: AlphaMesh) -> None:
    """Write m in the layout of Save_Setka_ALL_ALPHA, byte for byte."""
    out = [" ".join(str(int(h)) for h in m.header) + " "]
    out.append(str(len(m.px)))
    out += [f"{_g(x)} {_g(y)} {int(t)}" for x, y, t in zip(m.px, m.py, m.ptype)]
    for arr in (m.gran, m.gran_copy):
        out.append(str(len(arr)))
        out += [f"{a} {b} {t}" for a, b, t in arr]
    for fam in m.rails:
        out.append(str(len(fam)))
        for r in fam:
            out.append(" ".join(str(v) for v in r["m"]) + f" {_g(r['s'])} {r['type']}")
            out.append(str(len(r["all"])))
            out += [str(v) for v in r["all"]]
            out.append(str(len(r["key"])))
            out += [str(v) for v in r["key"]]
    out.append(str(len(m.cell_type)))
    for t, c, g in zip(m.cell_type, m.contour, m.cell_grans):
        out.append(str(int(t)))
        out.append(str(len(c)))
        out += [str(v) for v in c]
        out.append(str(len(g)))
        out += [f"{a} {b}" for a, b in g]
    for arr in (m.gran_links, m.copy_links):
        out.append(str(len(arr)))
        out += [f"{a} {b} {c}" for a, b, c in arr]
    out.append(str(len(m.point_cells)))
    for pc in m.point_cells:
        out.append(str(len(pc)))
        out += [str(v) for v in pc]
    for line in m.lines:
        out.append(str(len(line)))
        out += [str(v) for v in line]
    out.append(str(len(m.state)))
    for row in m.state:
        vals = [_g(v) for v in row]
        out.append(" ".join(vals[0:5]))
        for k in range(4):
            out.append(" ".join(vals[5 + 4 * k: 9 + 4 * k]))
        out.append(" ".join(vals[21:24]))
        for j in range(4):
            out.append(" ".join(vals[24 + 4 * j: 28 + 4 * j]))
        out.append(" ".join(vals[40:43]))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def alpha_from_polar(g: PolarGrid, prim, src=None, header=None) -> AlphaMesh:
    """A polar grid and its state as the reference holds them: the faces created per
    cell (radial face above, polar face towards j+1, each main face with its copy; the
    outer and upper-axis boundaries in that pass), then the lower axis and the inner
    sphere; cells with faces [left, above, right, below] (Cell::Grans order); the
    cells of every point in cell order.  prim: 5 or 21 fields; src: the 19 MK1 inputs
    (H_n[4], H_u[4], H_v[4], H_T[4], k_u, k_v, k_T) or None."""
    ni, nj = g.ni, g.nj
    nc = ni * nj
    nid = lambda i, j: i * (nj + 1) + j          # noqa: E731
    cid = lambda i, j: i * nj + j                # noqa: E731
    gran, gcopy, glink, clink = [], [], [], []
    own1, own2 = [None] * nc, [None] * nc        # (main number, copy number)

    def face(a, b, t, master, sosed):
        k = len(gran)
        gran.append((a, b, t))
        gcopy.append((b, a, t))
        glink.append((master, sosed if sosed is not None else -1, k))
        clink.append((sosed if sosed is not None else -1, master, k))
        return k

    for i in range(ni):
        for j in range(nj):
            c = cid(i, j)
            c1, c2, c3 = nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if i + 1 < ni:
                own1[c] = face(c1, c2, GEL_F_USUALY, c, cid(i + 1, j))
            else:
                own1[c] = face(c1, c2, int(g.outer_type[j]), c, None)
            if j + 1 < nj:
                own2[c] = face(c2, c3, GEL_F_USUALY, c, cid(i, j + 1))
            else:
                own2[c] = face(c2, c3, GEL_F_AXIS, c, None)
    # boundary copies have no master (the copy's master is the neighbour, absent)
    for k, (a, b, t) in enumerate(glink):
        if b == -1:
            clink[k] = (-1, a, k)
    cell_grans = []
    for i in range(ni):
        for j in range(nj):
            c = cid(i, j)
            if j > 0:
                e0 = (own2[cid(i, j - 1)], 0)
            else:
                k = face(nid(i, 0), nid(i + 1, 0), GEL_F_AXIS, c, None)
                clink[k] = (-1, c, k)
                e0 = (k, 1)
            if i > 0:
                e3 = (own1[cid(i - 1, j)], 0)
            else:
                k = face(nid(0, j + 1), nid(0, j), GEL_F_INNER_SPHERE, c, None)
                clink[k] = (-1, c, k)
                e3 = (k, 1)
            cell_grans.append([e0, (own1[c], 1), (own2[c], 1), e3])
    contour = [[nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)] for i in range(ni) for j in range(nj)]
    pc = [[] for _ in range(g.nnodes)]
    for c, q in enumerate(contour):
        for p in q:
            pc[p].append(c)
    prim = np.asarray(prim, np.float64).reshape(-1, nc)
    state = np.zeros((nc, len(STATE_FIELDS)))
    state[:, :prim.shape[0]] = prim.T
    if src is not None:
        s = np.asarray(src, np.float64).reshape(19, nc)
        for j in range(4):
            state[:, 24 + 4 * j] = s[j]
            state[:, 25 + 4 * j] = s[4 + j]
            state[:, 26 + 4 * j] = s[8 + j]
            state[:, 27 + 4 * j] = s[12 + j]
        state[:, 40:43] = s[16:19].T
    return AlphaMesh(list(header) if header is not None else [0] * 8, g.x.copy(), g.y.copy(),
                     np.full(g.nnodes, P_NO, np.int64), np.array(gran, np.int64), np.array(gcopy, np.int64),
                     [[], [], [], []], g.cell_type.astype(np.int64), contour, cell_grans, np.array(glink, np.int64),
                     np.array(clink, np.int64), pc, [[], [], []], state)


P_NO = 11      # Point_type::P_No (Point.h:7-22), the type of a plain grid point


def polar_from_alpha(m: AlphaMesh):
    """(PolarGrid, prim[21][nc], src[19][nc]) of a checkpoint whose cells form a
    logically structured polar grid numbered as alpha_from_polar numbers it; raises
    ValueError otherwise."""
    nc = len(m.cell_type)
    c0 = m.contour[0]
    if len(c0) != 4:
        raise ValueError("cells are not quadrilaterals")
    nj = c0[1] - c0[0] - 1
    if nj < 1 or nc % nj:
        raise ValueError("not a polar grid")
    ni = nc // nj
    if len(m.px) != (ni + 1) * (nj + 1):
        raise ValueError("not a polar grid")
    for c, q in enumerate(m.contour):
        i, j = divmod(c, nj)
        want = [i * (nj + 1) + j, (i + 1) * (nj + 1) + j, (i + 1) * (nj + 1) + j + 1, i * (nj + 1) + j + 1]
        if list(q) != want:
            raise ValueError(f"cell {c} is not cell ({i}, {j}) of a polar grid")
    outer = np.empty(nj, np.int8)
    for j in range(nj):
        k, main = m.cell_grans[(ni - 1) * nj + j][1]
        outer[j] = m.gran[k][2] if main else m.gran_copy[k][2]
    g = PolarGrid(ni, nj, m.px, m.py, m.cell_type.astype(np.int8), outer)
    prim = np.ascontiguousarray(m.state[:, :21].T)
    src = np.empty((19, nc))
    for j in range(4):
        src[j] = m.state[:, 24 + 4 * j]
        src[4 + j] = m.state[:, 25 + 4 * j]
        src[8 + j] = m.state[:, 26 + 4 * j]
        src[12 + j] = m.state[:, 27 + 4 * j]
    src[16:19] = m.state[:, 40:43].T
    return g, prim, np.ascontiguousarray(src)
