"""Write the golden fixtures of tests/golden/ by running the compiled reference.

Runs only in the build container, where /root/reference exists: oracle/Makefile
compiles the reference sources plus oracle/ref_harness.cpp into oracle/_ref/, and this
script feeds the harness seeded inputs and stores inputs and outputs as .npz files.
The reference's sweep has races on its shared time step (Setka.cpp:10141,
:10388-10393), so the harness runs with OMP_NUM_THREADS=1.

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py godunov    # one family (godunov, track, mf, inner, checkpoint, block,
                                                  # blocktrack, mc)
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gelios_2d_amd.grid import polar_grid  # noqa: E402
from gelios_2d_amd.params import (GEL_C_1, GEL_C_2, GEL_F_EXTERN, GEL_F_INPUT,  # noqa: E402
                                  reference_params)
from oracle import oracle as ora  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def harness(mode, blob: bytes, tmp):
    ora.build()
    fin = os.path.join(tmp, f"{mode}.in")
    fout = os.path.join(tmp, f"{mode}.out")
    with open(fin, "wb") as f:
        f.write(blob)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([ora.REF_HARNESS, mode, fin, fout], check=True, cwd=tmp, env=env,
                   stdout=subprocess.DEVNULL)
    with open(fout, "rb") as f:
        return f.read()


def gen_hllc(rng, tmp, n=2000):
    roL = rng.uniform(0.05, 3.0, n)
    roR = rng.uniform(0.05, 3.0, n)
    pL = rng.uniform(0.01, 3.0, n)
    pR = rng.uniform(0.01, 3.0, n)
    QL = rng.uniform(0.0, 5.0, n)
    QR = rng.uniform(0.0, 5.0, n)
    # a quarter of the faces carry a strong common drift (supersonic branches)
    drift = np.where(rng.random(n) < 0.25, rng.normal(0, 6.0, (2, n)), 0.0)
    uL, vL = rng.normal(0, 1.0, (2, n)) + drift
    uR, vR = rng.normal(0, 1.0, (2, n)) + drift
    ang = rng.uniform(0, 2 * np.pi, n)
    n1, n2 = np.cos(ang), np.sin(ang)
    W = np.where(rng.random(n) < 0.5, 0.0, rng.normal(0, 1.5, n))
    rad = rng.uniform(0.01, 1.0, n)
    metod = (np.arange(n) % 2).astype(np.float64)
    nul = ((np.arange(n) // 2) % 2).astype(np.float64)
    inp = np.stack([roL, QL, pL, uL, vL, roR, QR, pR, uR, vR, W, n1, n2, rad, metod, nul], axis=1)
    out = harness("hllc", np.int32(n).tobytes() + inp.tobytes(), tmp)
    res = np.frombuffer(out, np.float64).reshape(n, 9)
    np.savez_compressed(os.path.join(OUT, "hllc_vectors.npz"), inp=inp, out=res)


def gen_helpers(rng, tmp, n=1000):
    x1, x3 = rng.normal(0, 1, (2, n))
    x2 = x1 + rng.uniform(0.1, 1.0, n) * np.sign(rng.normal(size=n))
    t1, t2, t3 = rng.normal(0, 1, (3, n))
    # exact ties and zeros exercise the sign/minmod branches
    t2[::7] = t1[::7]
    t3[::11] = t2[::11]
    x1[::13] = 0.0
    y = rng.normal(0, 1, n)
    inp = np.stack([x1, t1, x2, t2, x3, t3, y], axis=1)
    out = harness("helpers", np.int32(n).tobytes() + inp.tobytes(), tmp)
    res = np.frombuffer(out, np.float64).reshape(n, 8)
    np.savez_compressed(os.path.join(OUT, "helpers_vectors.npz"), inp=inp, out=res)


def gen_cell(rng, tmp, n=300):
    rows = []
    for k in range(n):
        # perturbed polar quads; centres at x < 0 or r > 1500/RR_ keep Get_Sourse_MK1 on
        # its own-cell branch (Cell.cpp:590-595)
        r0 = rng.uniform(4.0, 20.0)
        th0 = rng.uniform(0.05, np.pi - 0.3)
        dr, dth = rng.uniform(0.05, 1.0), rng.uniform(0.02, 0.25)
        pts = [(r0, th0), (r0 + dr, th0), (r0 + dr, th0 + dth), (r0, th0 + dth)]
        q = []
        for (r, t) in pts:
            r *= 1 + 0.02 * rng.normal()
            q += [r * np.cos(t), r * np.sin(t)]
        uvrp = [rng.normal(), rng.normal(), rng.uniform(0.1, 3), rng.uniform(0.1, 3)]
        fld = list(rng.uniform(0.01, 1.0, 4)) + list(rng.normal(0, 2, 4)) + list(rng.normal(0, 2, 4)) + \
            list(rng.uniform(0.1, 3.0, 4)) + list(rng.uniform(0.33, 3.0, 3))
        if k % 10 == 0:
            fld[k % 4] = 0.0      # population with no density: clamp branch
        rows.append(q + uvrp + fld)
    inp = np.array(rows, np.float64)
    out = harness("cell", np.int32(n).tobytes() + inp.tobytes(), tmp)
    res = np.frombuffer(out, np.float64).reshape(n, 7)
    np.savez_compressed(os.path.join(OUT, "cell_vectors.npz"), inp=inp, out=res)


def gen_hlld(rng, tmp, n=1200):
    """Solvers::HLLDQ_Korolkov, metod 3, on the normals it supports (axis-aligned,
    Solvers.cpp:1664-1697) and a face at rest (:1643)."""
    roL, roR = rng.uniform(0.05, 3.0, (2, n))
    pL, pR = rng.uniform(0.01, 3.0, (2, n))
    QL, QR = rng.uniform(0.0, 5.0, (2, n))
    drift = np.where(rng.random(n) < 0.25, rng.normal(0, 5.0, (3, n)), 0.0)
    vL = rng.normal(0, 1.0, (3, n)) + drift
    vR = rng.normal(0, 1.0, (3, n)) + drift
    bscale = np.where(rng.random(n) < 0.2, 0.0, rng.uniform(0.1, 4.0, n))   # some B = 0 faces
    BL = rng.normal(0, 1.0, (3, n)) * bscale
    BR = rng.normal(0, 1.0, (3, n)) * bscale
    # a common normal field on half of the faces (div B = 0 across the face)
    k = np.arange(n)
    nsel = k % 4
    nrm = np.array([[1.0, 0, 0], [-1.0, 0, 0], [0, 1.0, 0], [0, -1.0, 0]])[nsel]
    same = rng.random(n) < 0.5
    for i in range(n):
        if same[i]:
            a = int(np.argmax(np.abs(nrm[i])))
            BR[a, i] = BL[a, i]
    inp = np.concatenate([roL[:, None], QL[:, None], pL[:, None], vL.T, BL.T,
                          roR[:, None], QR[:, None], pR[:, None], vR.T, BR.T, nrm, np.full((n, 1), 3.0)], axis=1)
    out = harness("hlld", np.int32(n).tobytes() + np.ascontiguousarray(inp).tobytes(), tmp)
    res = np.frombuffer(out, np.float64).reshape(n, 10)
    np.savez_compressed(os.path.join(OUT, "hlld_vectors.npz"), inp=inp, out=res)


def gen_godunov(rng, tmp, n=1600):
    """Solvers::Godunov_Solver_Alexashov (gamma 5/3, kurant 0.1): states drawn so that the
    four wave patterns (two shocks, shock + fan, two fans, vacuum) all occur, with
    pI > pII on about half of the faces (the swap of Solvers.cpp:750-777), face speeds
    W != 0 on half, an out-of-plane velocity on a quarter, in-plane unit normals."""
    ro = rng.uniform(0.05, 3.0, (2, n))
    p = rng.uniform(0.01, 3.0, (2, n))
    c = np.sqrt(5.0 / 3.0 * p / ro)
    ang = rng.uniform(0, 2 * np.pi, n)
    nrm = np.stack([np.cos(ang), np.sin(ang), np.zeros(n)], axis=1)
    # normal velocity jump enL - enR in units of the sound speeds: from strongly
    # separating (vacuum, < -3 (cL + cR)) to strongly colliding (two shocks)
    jump = rng.uniform(-5.0, 4.0, n) * (c[0] + c[1])
    tang = rng.normal(0, 1.0, (2, n))
    drift = rng.normal(0, 1.0, n)
    enL, enR = drift + 0.5 * jump, drift - 0.5 * jump
    vL = np.stack([enL * nrm[:, 0] + tang[0] * nrm[:, 1], enL * nrm[:, 1] - tang[0] * nrm[:, 0]])
    vR = np.stack([enR * nrm[:, 0] + tang[1] * nrm[:, 1], enR * nrm[:, 1] - tang[1] * nrm[:, 0]])
    wz = np.where(rng.random((2, n)) < 0.25, rng.normal(0, 1.0, (2, n)), 0.0)
    W = np.where(rng.random(n) < 0.5, 0.0, rng.normal(0, 1.5, n) * (c[0] + c[1]) / 2)
    dist = rng.uniform(0.01, 1.0, n)
    inp = np.concatenate([ro[0][:, None], p[0][:, None], vL.T, wz[0][:, None],
                          ro[1][:, None], p[1][:, None], vR.T, wz[1][:, None], nrm, W[:, None], dist[:, None]], axis=1)
    out = harness("godunov", np.int32(n).tobytes() + np.ascontiguousarray(inp).tobytes(), tmp)
    res = np.frombuffer(out, np.float64).reshape(n, 9)
    np.savez_compressed(os.path.join(OUT, "godunov_vectors.npz"), inp=inp, out=res)


def harness_grid(ni=16, nj=10, n_c1=3):
    g = polar_grid(ni, nj, 4.0, 12.0, spacing="log")
    ct = np.full(ni * nj, GEL_C_2, np.int8)
    ct[: n_c1 * nj] = GEL_C_1
    g.cell_type = ct
    return g


def initial_state(g, rng, prm):
    """Wind from the inner sphere (the Inner_sphere ghost formula, Gran.cpp:234-265)
    blended into the reference's Input state, with a seeded 1% density noise."""
    cx, cy = g.cell_centers()
    r = np.hypot(cx, cy)
    ro_w = 1.0 / (prm.in_rho_div * r * r)
    p_w = ro_w * prm.in_p_a * prm.in_p_a / prm.in_p_den
    u_w, v_w = prm.in_u * cx / r, prm.in_u * cy / r
    s = 0.5 * (1 - np.tanh((r - 7.0) / 1.5))
    ro = s * ro_w + (1 - s) * prm.inf_rho
    p = s * p_w + (1 - s) * prm.inf_p
    u = s * u_w + (1 - s) * prm.inf_u
    v = s * v_w + (1 - s) * prm.inf_v
    Q = s * ro_w + (1 - s) * prm.inf_q
    ro = ro * (1 + 0.01 * rng.uniform(-1, 1, ro.shape))
    return np.stack([ro, p, u, v, Q])


def sweep_blob(g, steps, prim, traj=None, src=None):
    h = np.array([g.ni, g.nj, steps, 1 if traj is not None else 0, 1 if src is not None else 0], np.int32)
    parts = [h.tobytes(), g.x.tobytes(), g.y.tobytes()]
    if traj is not None:
        parts.append(np.ascontiguousarray(traj, np.float64).tobytes())
    parts += [g.cell_type.astype(np.int32).tobytes(), g.outer_type.astype(np.int32).tobytes(),
              np.ascontiguousarray(prim, np.float64).tobytes()]
    if src is not None:
        parts.append(np.ascontiguousarray(src, np.float64).tobytes())
    return b"".join(parts)


def parse_sweep(out, g):
    nc = g.ncells
    o = 0
    prim = np.frombuffer(out, np.float64, 5 * nc, o).reshape(5, nc); o += 5 * nc * 8
    contour = np.frombuffer(out, np.int32, 4 * nc, o).reshape(nc, 4); o += 4 * nc * 4
    nface = np.frombuffer(out, np.int32, nc, o); o += nc * 4
    slots = np.frombuffer(out, np.int32, nc * 8 * 6, o).reshape(nc, 8, 6)
    return prim.copy(), contour.copy(), nface.copy(), slots.copy()


def gen_sweeps(rng, tmp):
    prm = reference_params()
    # static grid, 40 steps of Go_5_komponent__MK2
    g = harness_grid()
    prim0 = initial_state(g, rng, prm)
    steps = 40
    out = harness("sweep", sweep_blob(g, steps, prim0), tmp)
    prim, contour, nface, slots = parse_sweep(out, g)
    np.savez_compressed(os.path.join(OUT, "sweep_static.npz"), ni=g.ni, nj=g.nj, x=g.x, y=g.y,
                        cell_type=g.cell_type, outer_type=g.outer_type, steps=steps, prim0=prim0,
                        prim=prim, contour=contour, nface=nface, slots=slots)

    # moving grid: every step a seeded radial displacement of the nodes
    steps = 6
    traj = []
    x, y = g.x.copy(), g.y.copy()
    for s in range(steps):
        f = 1.0 + 1e-9 * rng.uniform(-1, 1, x.shape)
        traj.append(np.stack([x * f, y * f]))
        x, y = x * f, y * f
    traj = np.array(traj)
    prim0 = initial_state(g, rng, prm)
    out = harness("sweep", sweep_blob(g, steps, prim0, traj=traj), tmp)
    prim, contour, nface, slots = parse_sweep(out, g)
    np.savez_compressed(os.path.join(OUT, "sweep_moving.npz"), ni=g.ni, nj=g.nj, x=g.x, y=g.y,
                        cell_type=g.cell_type, outer_type=g.outer_type, steps=steps, prim0=prim0,
                        traj=traj, prim=prim, contour=contour, nface=nface, slots=slots)

    # static grid with charge-exchange sources from 4 populations
    steps = 25
    nc = g.ncells
    src = np.concatenate([rng.uniform(0.05, 0.5, (4, nc)), rng.normal(-2.0, 0.5, (4, nc)),
                          rng.normal(0.0, 0.5, (4, nc)), rng.uniform(0.5, 2.0, (4, nc)),
                          rng.uniform(0.5, 2.0, (3, nc))])
    prim0 = initial_state(g, rng, prm)
    out = harness("sweep", sweep_blob(g, steps, prim0, src=src), tmp)
    prim, contour, nface, slots = parse_sweep(out, g)
    np.savez_compressed(os.path.join(OUT, "sweep_sources.npz"), ni=g.ni, nj=g.nj, x=g.x, y=g.y,
                        cell_type=g.cell_type, outer_type=g.outer_type, steps=steps, prim0=prim0,
                        src=src, prim=prim, contour=contour, nface=nface, slots=slots)


def track_blob(c, steps, src=None):
    g, t = c.grid, c.tracking
    h = np.array([g.ni, g.nj, steps, 1 if src is not None else 0, t.row_inner_shock, t.row_contact,
                  t.row_outer_shock, min(t.outer_faces, g.nj)], np.int32)
    parts = [h.tobytes(), g.x.tobytes(), g.y.tobytes(), np.ascontiguousarray(c.rails, np.float64).tobytes(),
             g.cell_type.astype(np.int32).tobytes(), g.outer_type.astype(np.int32).tobytes(),
             np.ascontiguousarray(c.prim0, np.float64).tobytes()]
    if src is not None:
        parts.append(np.ascontiguousarray(src, np.float64).tobytes())
    return b"".join(parts)


def gen_tracked(tmp):
    """Fitted-surface tracking: Go_5_komponent__MK2 with Move_surface and the A rails
    of Move_Setka_Calculate_2 on tracked heliospheres (the reference's sweep runs HLL,
    met = 0 at Setka.cpp:10210): 39 x 10 cells for 30 steps with the outer shock moved
    by its first 7 faces, 44 x 16 cells for 60 steps with all of them, and 39 x 12 cells
    for 40 steps with a spherical inner shock, where the radial shocked wind has no
    tangential speed and the inner shock's faces move both of their nodes."""
    from gelios_2d_amd.configs import tracked_heliosphere, tracked_spherical_ts
    cases = {"a": (10, (24, 5, 5, 6), 30, 31, 7, None), "b": (16, (26, 6, 6, 7), 60, 32, 16, None),
             "c": (12, (24, 5, 5, 6), 40, 33, 12, tracked_spherical_ts)}
    res = {}
    for tag, (nj, blocks, steps, seed, outer, kr) in cases.items():
        c = tracked_heliosphere(nj, blocks, steps, seed, flux=0, outer_faces=outer, key_radii=kr)
        out = harness("track", track_blob(c, steps), tmp)
        g = c.grid
        nc, nn = g.ncells, g.nnodes
        prim = np.frombuffer(out, np.float64, 5 * nc, 0).reshape(5, nc).copy()
        x = np.frombuffer(out, np.float64, nn, 5 * nc * 8).copy()
        y = np.frombuffer(out, np.float64, nn, (5 * nc + nn) * 8).copy()
        res[tag] = dict(nj=nj, blocks=np.array(blocks), seed=seed, steps=steps, outer_faces=outer,
                        spherical_ts=kr is not None, prim0=c.prim0,
                        prim=prim, x=x, y=y, x0=g.x, y0=g.y, rails=c.rails, cell_type=g.cell_type,
                        outer_type=g.outer_type)
    np.savez_compressed(os.path.join(OUT, "sweep_tracked.npz"),
                        **{f"{k}_{tag}": v for tag, d in res.items() for k, v in d.items()})


def mf_state(g, rng, prm):
    """21 fields: the plasma of initial_state and four hydrogen fluids (seeded, positive
    densities and pressures, flows around the inflow speed)."""
    plasma = initial_state(g, rng, prm)
    nc = g.ncells
    h = []
    for k in range(4):
        h += [rng.uniform(0.05, 1.0, nc), rng.uniform(0.02, 0.6, nc),
              prm.inf_u * rng.uniform(0.3, 1.2, nc), rng.normal(0.0, 0.4, nc)]
    return np.concatenate([plasma, np.array(h)])


def mf_blob(g, steps, prim, traj=None):
    h = np.array([g.ni, g.nj, steps, 1 if traj is not None else 0], np.int32)
    parts = [h.tobytes(), g.x.tobytes(), g.y.tobytes()]
    if traj is not None:
        parts.append(np.ascontiguousarray(traj, np.float64).tobytes())
    parts += [g.cell_type.astype(np.int32).tobytes(), g.outer_type.astype(np.int32).tobytes(),
              np.ascontiguousarray(prim, np.float64).tobytes()]
    return b"".join(parts)


def gen_multifluid(rng, tmp):
    """Setka::Go_5_komponent_2 (plasma + four hydrogen fluids, HLLC, charge exchange) on
    the 16 x 10 harness grid with every source region: C_1 wind rows, C_2, C_3, C_4 and
    C_5 rows, two frozen C_centr cells; 30 static steps, and 6 steps with the nodes
    moved by seeded radial displacements."""
    from gelios_2d_amd.params import GEL_C_3, GEL_C_4, GEL_C_5, GEL_C_CENTR, GEL_PHYS_GAS4H
    prm = reference_params()
    g = harness_grid()
    ct = g.cell_type.copy()
    nj = g.nj
    ct[6 * nj: 8 * nj] = GEL_C_3
    ct[8 * nj: 11 * nj] = GEL_C_4
    ct[13 * nj:] = GEL_C_5
    ct[4 * nj + 3] = GEL_C_CENTR
    ct[9 * nj + 6] = GEL_C_CENTR
    g.cell_type = ct
    steps = 30
    prim0 = mf_state(g, rng, prm)
    out = harness("mf", mf_blob(g, steps, prim0), tmp)
    prim = np.frombuffer(out, np.float64).reshape(21, g.ncells).copy()
    res = dict(ni=g.ni, nj=g.nj, x=g.x, y=g.y, cell_type=g.cell_type, outer_type=g.outer_type,
               static_steps=steps, static_prim0=prim0, static_prim=prim)
    steps = 6
    traj = []
    x, y = g.x.copy(), g.y.copy()
    for _ in range(steps):
        f = 1.0 + 1e-9 * rng.uniform(-1, 1, x.shape)
        traj.append(np.stack([x * f, y * f]))
        x, y = x * f, y * f
    traj = np.array(traj)
    prim0 = mf_state(g, rng, prm)
    out = harness("mf", mf_blob(g, steps, prim0, traj), tmp)
    prim = np.frombuffer(out, np.float64).reshape(21, g.ncells).copy()
    res.update(moving_steps=steps, moving_prim0=prim0, moving_prim=prim, traj=traj)
    np.savez_compressed(os.path.join(OUT, "sweep_multifluid.npz"), **res)
    assert GEL_PHYS_GAS4H == 2


def gen_inner(tmp):
    """Setka::Go_stationary_5_komponent_inner_MK2 on the inner region, 24 x 12 cells
    (C_1 rows, C_2 rows, two C_centr cells), 40 steps without sources and 25 with the
    MK1 closure of seeded hydrogen moments."""
    from gelios_2d_amd.configs import inner_region
    res = {}
    for tag, steps, seed, srcs in (("plain", 40, 41, False), ("sources", 25, 42, True)):
        w = inner_region(24, 12, steps, seed, sources=srcs)
        g = w.grid
        h = np.array([g.ni, g.nj, steps, 1 if srcs else 0], np.int32)
        parts = [h.tobytes(), g.x.tobytes(), g.y.tobytes(), g.cell_type.astype(np.int32).tobytes(),
                 g.outer_type.astype(np.int32).tobytes(), np.ascontiguousarray(w.prim0).tobytes()]
        if srcs:
            parts.append(w.src.tobytes())
        out = harness("inner", b"".join(parts), tmp)
        res[f"{tag}_prim0"] = w.prim0
        res[f"{tag}_prim"] = np.frombuffer(out, np.float64).reshape(5, g.ncells).copy()
        res[f"{tag}_steps"] = steps
        res[f"{tag}_seed"] = seed
        if srcs:
            res[f"{tag}_src"] = w.src
    res.update(ni=24, nj=12)
    np.savez_compressed(os.path.join(OUT, "sweep_inner.npz"), **res)


def gen_checkpoint(rng, tmp):
    """Setka::Save_Setka_ALL_ALPHA of a 7 x 5 polar grid (C_1 and C_2 rows) with all 21
    fields and the 19 MK1 inputs, and Download_Setka_ALL_ALPHA_2_0 + Save of that file
    (the reference reads its own checkpoint back unchanged).  The file is kept as bytes."""
    prm = reference_params()
    g = harness_grid(7, 5, 2)
    nc = g.ncells
    prim = mf_state(g, rng, prm)
    src = np.concatenate([rng.uniform(0.05, 0.5, (4, nc)), rng.normal(-2.0, 0.5, (4, nc)),
                          rng.normal(0.0, 0.5, (4, nc)), rng.uniform(0.5, 2.0, (4, nc)),
                          rng.uniform(0.5, 2.0, (3, nc))])
    h = np.array([g.ni, g.nj, 1], np.int32)
    blob = b"".join([h.tobytes(), g.x.tobytes(), g.y.tobytes(), g.cell_type.astype(np.int32).tobytes(),
                     g.outer_type.astype(np.int32).tobytes(), prim.tobytes(), src.tobytes()])
    text = harness("save", blob, tmp)
    np.savez_compressed(os.path.join(OUT, "checkpoint_polar.npz"), ni=g.ni, nj=g.nj, x=g.x, y=g.y,
                        cell_type=g.cell_type, outer_type=g.outer_type, prim=prim, src=src,
                        ref_text=np.frombuffer(text, np.uint8))


BLOCK_GRIDS = {"a": ((4, 3, 3, 10, 22, 6, 7, 7), 30), "b": ((6, 3, 4, 8, 24, 6, 8, 8), 40)}


def parse_block(b: bytes) -> dict:
    """The harness's block-mode output (oracle/ref_harness.cpp, mode_block)."""
    nn, nc = np.frombuffer(b, np.int32, 2)
    off = 8

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(b, dt, n, off).copy()
        off += a.nbytes
        return a
    d = {"x": take(np.float64, nn), "y": take(np.float64, nn), "ptype": take(np.int32, nn)}
    d["ncorner"] = take(np.int32, nc)
    d["corners"] = take(np.int32, int(d["ncorner"].sum()))
    d["cell_type"] = take(np.int32, nc)
    d["nface"] = take(np.int32, nc)
    d["slots"] = take(np.int32, nc * 8 * 6).reshape(nc, 8, 6)
    d["prim0"] = take(np.float64, 5 * nc).reshape(5, nc)
    d["prim"] = take(np.float64, 5 * nc).reshape(5, nc)
    d["x_after"] = take(np.float64, nn)
    d["y_after"] = take(np.float64, nn)
    assert off == len(b)
    return d


def gen_block(tmp):
    """The reference's own block grid: Setka(N1..M4) (Setka.cpp:19-1076) filled by
    Init_conditions, written by Save_Setka_ALL_ALPHA and read back by
    Download_Setka_ALL_ALPHA_2_0, then TVD_prepare and Go_5_komponent__MK2(steps) with
    stand-in fitted-surface lines (oracle/ref_harness.cpp, mode_block).  Kept: the
    checkpoint text, the face tables as the reference holds them, the initial state,
    the state and node positions after the steps."""
    ora.build()
    res = {}
    for tag, (N, steps) in BLOCK_GRIDS.items():
        fin, fout, ck = (os.path.join(tmp, f"block_{tag}.{e}") for e in ("in", "out", "txt"))
        with open(fin, "wb") as f:
            f.write(np.array(list(N) + [steps], np.int32).tobytes())
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([ora.REF_HARNESS, "block", fin, fout, ck], check=True, cwd=tmp, env=env,
                       stdout=subprocess.DEVNULL)
        with open(fout, "rb") as f:
            d = parse_block(f.read())
        with open(ck, "rb") as f:
            d["ref_text"] = np.frombuffer(f.read(), np.uint8)
        d["N"] = np.array(N, np.int32)
        d["steps"] = steps
        for k, v in d.items():
            res[f"{tag}_{k}"] = v
    np.savez_compressed(os.path.join(OUT, "sweep_block.npz"), **res)


def parse_block_tracked(b: bytes) -> dict:
    """The harness's blocktrack-mode output (oracle/ref_harness.cpp, mode_blocktrack)."""
    nn, nc = np.frombuffer(b, np.int32, 2)
    off = 8

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(b, dt, n, off).copy()
        off += a.nbytes
        return a
    d = {"x": take(np.float64, nn), "y": take(np.float64, nn), "ptype": take(np.int32, nn)}
    d["ncorner"] = take(np.int32, nc)
    d["corners"] = take(np.int32, int(d["ncorner"].sum()))
    d["cell_type"] = take(np.int32, nc)
    d["nface"] = take(np.int32, nc)
    d["slots"] = take(np.int32, nc * 8 * 6).reshape(nc, 8, 6)
    d["prim0"] = take(np.float64, 5 * nc).reshape(5, nc)
    rc = take(np.int32, 4)
    kind, m, s, alls, key = [], [], [], [], []
    for q in range(4):
        for _ in range(rc[q]):
            h = take(np.int32, 6)
            kind.append(q)
            m.append(h[:4])
            s.append(take(np.float64, 1)[0])
            alls.append(take(np.int32, h[4]))
            k = take(np.int32, h[5])
            key.append(np.concatenate([k, np.full(4 - len(k), -1, np.int32)]))
    d["rail_kind"] = np.array(kind, np.int32)
    d["rail_m"] = np.array(m, np.int32)
    d["rail_s"] = np.array(s, np.float64)
    d["rail_offset"] = np.concatenate([[0], np.cumsum([len(a) for a in alls])]).astype(np.int32)
    d["rail_all"] = np.concatenate(alls).astype(np.int32)
    d["rail_key"] = np.array(key, np.int32)
    lc = take(np.int32, 3)
    for q, name in enumerate(("contact", "inner", "outer")):
        d["line_" + name] = take(np.int32, 4 * lc[q]).reshape(lc[q], 4)
    for tag in ("1", "n"):
        d["prim_" + tag] = take(np.float64, 5 * nc).reshape(5, nc)
        d["x_" + tag] = take(np.float64, nn)
        d["y_" + tag] = take(np.float64, nn)
    assert off == len(b)
    return d


BLOCK_TRACKED = {"a": ((4, 3, 3, 10, 22, 6, 7, 7), 20), "b": ((6, 3, 4, 8, 24, 6, 8, 8), 20)}


def gen_block_tracked(tmp):
    """The reference's own block grid with its fitted surfaces moving: Setka(N1..M4) as
    the constructor builds it (rails of four kinds, Line_Contact / Line_Inner /
    Line_Outer), Init_conditions, TVD_prepare, then Go_5_komponent__MK2 for 1 step and
    for `steps` steps, each step running Move_surface and Move_Setka_Calculate_2
    (oracle/ref_harness.cpp, mode_blocktrack).  Kept: the grid, rails and lines, the
    initial state, and the state and node positions after both runs."""
    ora.build()
    res = {}
    for tag, (N, steps) in BLOCK_TRACKED.items():
        fin, fout = (os.path.join(tmp, f"blocktrack_{tag}.{e}") for e in ("in", "out"))
        with open(fin, "wb") as f:
            f.write(np.array(list(N) + [steps], np.int32).tobytes())
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([ora.REF_HARNESS, "blocktrack", fin, fout], check=True, cwd=tmp, env=env,
                       stdout=subprocess.DEVNULL)
        with open(fout, "rb") as f:
            d = parse_block_tracked(f.read())
        d["N"] = np.array(N, np.int32)
        d["steps"] = steps
        for k, v in d.items():
            res[f"{tag}_{k}"] = v
    np.savez_compressed(os.path.join(OUT, "sweep_block_tracked.npz"), **res)


def gen_mc(rng, tmp):
    """Point-wise pieces of the reference's Monte Carlo (oracle/ref_harness.cpp,
    mode_mcunit): Setka::Velosity_1..3 and MKmethod::int_1..3 on (u, cp) pairs, the
    random sensor (sensor.cpp) from seeds, and Change_Velosity4 with no splitting zones
    (I = 0) from seeded sensors, 50 samples per input."""
    x = np.arange(0.0, 7.0 + 1e-9, 0.05)
    dense = np.concatenate([np.stack([x * cp, np.full_like(x, cp)], 1) for cp in (0.5, 1.0, 2.5)])
    grid = np.array([(u, cp) for u in (0.0, 1e-6, 3e-6, 1e-3, 0.1, 0.5, 1.0, 2.0, 3.5, 5.0, 6.9, 8.0, 12.0, 20.0)
                     for cp in (0.2, 0.7, 1.0, 1.9, 4.0)])
    rates_in = np.concatenate([dense, grid]).astype(np.float64)
    nr = len(rates_in)
    seeds = np.array([[1, 0, 1], [12345, 678, 91], [32767, 32767, 1023], [2 * 4711 + 1, 29000, 555]], np.int32)
    nd = 400
    xin = []
    for X, cp in ((0.01, 1.0), (0.4, 0.3), (1.0, 1.0), (2.2, 0.8), (3.0, 2.0), (5.5, 0.5), (8.0, 1.0),
                  (12.0, 0.1), (0.8, 4.0), (1.7, 0.05), (6.9, 1.5), (0.05, 0.6)):
        U = rng.normal(size=3)
        d = rng.normal(size=3)
        V = U + X * d / np.linalg.norm(d)
        xin.append([*U, *V, cp])
    xin = np.array(xin, np.float64)
    nx, reps = len(xin), 50
    xseed = np.stack([2 * rng.integers(0, 16384, nx) + 1, rng.integers(0, 32768, nx), rng.integers(0, 1024, nx)],
                     1).astype(np.int32)
    blob = (np.int32(nr).tobytes() + rates_in.tobytes() + np.array([len(seeds), nd], np.int32).tobytes() +
            seeds.tobytes() + np.array([nx, reps], np.int32).tobytes() + xin.tobytes() + xseed.tobytes())
    out = harness("mcunit", blob, tmp)
    o = 0
    rates = np.frombuffer(out, np.float64, nr * 6, o).reshape(nr, 6)
    o += nr * 6 * 8
    draws = np.frombuffer(out, np.float64, len(seeds) * nd, o).reshape(len(seeds), nd)
    o += len(seeds) * nd * 8
    samples = np.frombuffer(out, np.float64, nx * reps * 4, o).reshape(nx, reps, 4)
    o += nx * reps * 4 * 8
    state = np.frombuffer(out, np.int32, nx * 3, o).reshape(nx, 3)
    np.savez_compressed(os.path.join(OUT, "mc_vectors.npz"), rates_in=rates_in, rates=rates, seeds=seeds,
                        draws=draws, xin=xin, xseed=xseed, samples=samples, state=state)


def main():
    only = sys.argv[1:]
    if only:
        with tempfile.TemporaryDirectory() as tmp:
            if "mc" in only:
                gen_mc(np.random.default_rng(20261208), tmp)
            if "godunov" in only:
                gen_godunov(np.random.default_rng(20261205), tmp)
            if "track" in only:
                gen_tracked(tmp)
            if "checkpoint" in only:
                gen_checkpoint(np.random.default_rng(20261207), tmp)
            if "inner" in only:
                gen_inner(tmp)
            if "mf" in only:
                gen_multifluid(np.random.default_rng(20261206), tmp)
            if "block" in only:
                gen_block(tmp)
            if "blocktrack" in only:
                gen_block_tracked(tmp)
        return
    rng = np.random.default_rng(20261203)
    with tempfile.TemporaryDirectory() as tmp:
        gen_hllc(rng, tmp)
        gen_helpers(rng, tmp)
        gen_cell(rng, tmp)
        gen_hlld(np.random.default_rng(20261204), tmp)
        gen_sweeps(rng, tmp)
        gen_godunov(np.random.default_rng(20261205), tmp)
        gen_tracked(tmp)
        gen_multifluid(np.random.default_rng(20261206), tmp)
        gen_inner(tmp)
        gen_checkpoint(np.random.default_rng(20261207), tmp)
        gen_block(tmp)
        gen_block_tracked(tmp)
        gen_mc(np.random.default_rng(20261208), tmp)
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
