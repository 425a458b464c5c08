"""The CPU oracle of the hydrogen Monte Carlo (oracle/mc_oracle.c) pinned against the
reference and against what the physics fixes.

- Against the compiled reference (tests/golden/mc_vectors.npz, oracle/ref_harness.cpp
  mode_mcunit): the random sensor (sensor.cpp:11-31), Setka::Velosity_1..3
  (Setka.cpp:15830-15866) and the exchange sampler MKmethod::Change_Velosity4 with I = 0
  (MKmethod.cpp:2270-2531) bit for bit; MKmethod::int_1..3 (MKmethod.cpp:2741-2858, the
  reference's polynomial fits) to the fits' accuracy.
- The rate integrals without the cross section's logarithm (a = 0) reduce to the
  reference's closed forms Velosity_1..3, which the tracer takes beyond u / cp = 7.
- The tracer (Setka::Fly_exchenge_Imit_Korol, Setka.cpp:12814-13981, in the device's
  form): in vacuum the density, bulk velocity and temperature of the inflowing Maxwellian
  hold in every cell (Liouville); in a plasma with the hydrogen's own Maxwellian the
  exchanges leave the total moments unchanged and carry no momentum or energy; the
  source estimated from the rate integrals (I) and from the sampled exchanges (II) agree.
"""
import os

import numpy as np
import pytest

from gelios_2d_amd.grid import polar_grid
from gelios_2d_amd.params import mc_reference_params
from oracle import oracle as ora

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mc_vectors.npz")
V_INF = -2.54338


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_sensor_matches_reference(gold):
    for seed, draws in zip(gold["seeds"], gold["draws"]):
        assert np.array_equal(ora.sensor_draws(seed, draws.size), draws)


def test_velosity_matches_reference(gold):
    r = ora.mc_rates(gold["rates_in"][:, 0], gold["rates_in"][:, 1])
    assert np.array_equal(r[:, :3], gold["rates"][:, :3])


def test_rate_integrals_match_reference_fits(gold):
    ri, ref = gold["rates_in"], gold["rates"]
    r = ora.mc_rates(ri[:, 0], ri[:, 1])
    m = ri[:, 0] / ri[:, 1] <= 7.0
    assert m.sum() > 400
    rel = np.abs(r[m, 3:] - ref[m, 3:]) / np.maximum(np.abs(ref[m, 3:]), 1e-12)
    assert rel.max() < 2e-5, rel.max(axis=0)
    # a sign or factor error in any of the three integrals fails here
    assert np.all(np.sign(r[m, 3:]) == np.sign(ref[m, 3:]))


def test_rate_integrals_reduce_to_closed_forms():
    """With sigma = 1 the three integrals are the reference's own closed forms:
    int_1 = Velosity_1, int_2 = -Velosity_2 / (cp^3 sqrtpi^3), int_3 = (3 cp^2 + 2 u^2) Velosity_3 / 2
    (the branch Fly_exchenge_Imit_Korol takes beyond u / cp = 7, Setka.cpp:13410-13420)."""
    u = np.array([1e-6, 0.01, 0.3, 1.0, 2.5, 4.0, 6.9, 0.5, 3.0])
    cp = np.array([1.0, 1.0, 0.7, 1.3, 0.9, 1.0, 1.0, 2.5, 0.6])
    V1, V2, V3, k1, k2, k3 = ora.mc_rates(u, cp, sig_a=0.0).T
    n3 = 1.77245385 ** 3
    np.testing.assert_allclose(k1, V1, rtol=1e-8)
    np.testing.assert_allclose(k2, -V2 / (cp ** 3 * n3), rtol=2e-6)
    np.testing.assert_allclose(k3, 0.5 * (3 * cp ** 2 + 2 * u ** 2) * V3, rtol=1e-8)


def test_exchange_sampler_matches_reference(gold):
    out, state = ora.mc_exchange(gold["xin"], gold["xseed"], gold["samples"].shape[1])
    assert np.array_equal(out, gold["samples"])
    assert np.array_equal(state, gold["state"])


def _run(prim_fn, n, seed=7, ni=8, nj=6):
    g = polar_grid(ni, nj, 1.0, 10.0)
    nc = ni * nj
    prim = np.zeros((5, nc))
    prim_fn(prim)
    out, st = ora.mc_run(g, prim, ora.mc_population(None, nc), mc_reference_params(n, seed=seed))
    return g, out, st


def _row_mean(g, field, vol):
    f = field.reshape(g.ni, g.nj)
    w = vol.reshape(g.ni, g.nj)
    return (f * w).sum(1) / w.sum(1)


def _volumes(g):
    x, y = g.x.reshape(g.ni + 1, g.nj + 1), g.y.reshape(g.ni + 1, g.nj + 1)
    v = np.zeros((g.ni, g.nj))
    for i in range(g.ni):
        for j in range(g.nj):
            px = [x[i, j], x[i + 1, j], x[i + 1, j + 1], x[i, j + 1]]
            py = [y[i, j], y[i + 1, j], y[i + 1, j + 1], y[i, j + 1]]
            A = G = 0.0
            for k in range(4):
                cr = px[k] * py[(k + 1) % 4] - py[k] * px[(k + 1) % 4]
                A += cr
                G += (py[k] + py[(k + 1) % 4]) * cr
            v[i, j] = abs(G / 6.0) * 2 * np.pi
    return v.ravel()


def test_vacuum_keeps_the_inflow_maxwellian():
    def vac(p):
        p[0] = 1e-200
        p[1] = 1e-200
    g, out, st = _run(vac, 60000)
    assert st["lost"] == 0 and st["primaries_in"] > 55000
    assert st["atoms"] == st["primaries_in"]          # no exchange, no secondary
    vol = _volumes(g)
    n, u, v, T = out[25], out[26], out[27], out[28]
    np.testing.assert_array_equal(out[3], n)          # all primaries (population 4)
    assert np.all(out[0:3] == 0.0)
    # row means (volume-weighted) to the statistics of 6e4 atoms; every cell within 15 %
    assert np.abs(_row_mean(g, n, vol) - 1.0).max() < 0.03
    assert np.abs(n - 1.0).max() < 0.15
    assert np.abs(_row_mean(g, u, vol) - V_INF).max() < 0.02
    assert np.abs(_row_mean(g, v, vol)).max() < 0.03
    assert np.abs(_row_mean(g, T, vol) - 1.0).max() < 0.04
    assert abs(np.average(n, weights=vol) - 1.0) < 0.01
    assert np.all(out[19:25] == 0.0)                  # no exchange, no source


def test_equilibrium_plasma_leaves_the_moments():
    """The plasma with the hydrogen's Maxwellian (velocity v_inf along the axis, cp = 1):
    detailed balance holds for any cross section, so charge exchange changes neither the
    total density, velocity and temperature nor carries momentum or energy."""
    def eq(p):
        p[0] = 0.5
        p[1] = 0.5
        p[2] = V_INF
    g, out, st = _run(eq, 60000)
    assert st["lost"] == 0 and st["atoms"] > 2 * st["primaries_in"]
    vol = _volumes(g)
    n, u, T = out[25], out[26], out[28]
    assert np.abs(_row_mean(g, n, vol)[2:] - 1.0).max() < 0.03
    assert abs(np.average(n, weights=vol) - 1.0) < 0.015
    assert np.abs(_row_mean(g, u, vol) - V_INF).max() < 0.02
    assert np.abs(_row_mean(g, T, vol) - 1.0).max() < 0.04
    # secondaries are population 2 (C_2 cells); the primaries decay inward
    assert np.all(out[1] > 0.0) and np.all(out[0] == 0.0)
    hn = _row_mean(g, out[3], vol)
    assert hn[-1] > hn[0]
    # the momentum and energy sources vanish against those of the same plasma at rest
    rest = _run(lambda p: (p.__setitem__(0, 0.5), p.__setitem__(1, 0.5)), 20000)[1]
    ref = np.abs(rest[19:22].sum(1))
    I = np.abs(out[19:22].sum(1))
    assert I[0] < 0.02 * ref[0] and I[1] < 0.02 * ref[0] and I[2] < 0.02 * ref[2], (I, ref)


def test_rate_and_sampled_sources_agree():
    """I (MKmethod::int_2 / int_3 at the exchange point, Setka.cpp:13420-13432) and II (the
    sampled exchanges, :13466-13470) estimate the same momentum and energy transfer."""
    def hot(p):
        p[0] = 0.3
        p[1] = 0.6
        p[2] = 0.5
        p[3] = 0.2
    g, out, st = _run(hot, 100000, seed=11)
    assert st["lost"] == 0
    I, II = out[19:22].sum(1), out[22:25].sum(1)
    assert abs(I[0] - II[0]) < 0.02 * abs(I[0])
    assert abs(I[2] - II[2]) < 0.05 * abs(I[2])
    assert abs(I[1] - II[1]) < 0.15 * abs(I[1])
    # momentum flows from the plasma to the atoms along the relative drift
    assert I[0] < 0 and I[2] > 0
    k = out[16:19]
    assert np.all((k >= 0.33) & (k <= 3.0))
    assert np.all(np.isfinite(out))


# ------------------------------------------------- the walk on a general mesh ----

def _hot(nc):
    p = np.zeros((5, nc))
    p[0], p[1], p[2], p[3] = 0.3, 0.6, 0.5, 0.2
    return p


def test_mesh_walk_equals_polar_walk():
    """The polar grid given as a general mesh (face slots in Cell::Grans order, every face's
    nodes as its cell sees them): the walk over each cell's slots, with Axis slots skipped,
    a Sosed followed, the inner sphere crossed and the outer faces left, traces the same
    atoms as the polar walk -- equal counts, the tallies to the last bits of crossings
    solved from the other end of a face."""
    from gelios_2d_amd.mesh import mesh_from_polar
    g = polar_grid(8, 6, 1.0, 10.0)
    m = mesh_from_polar(g)
    prim = _hot(g.ncells)
    pop = ora.mc_population(g.cell_type, g.ncells)
    p = mc_reference_params(20000, seed=5)
    p.max_segments = 100_000
    out_p, st_p = ora.mc_run(g, prim, pop, p)
    out_m, st_m = ora.mc_run_mesh(m, prim, pop, p)
    for k in ("primaries_in", "atoms", "segments", "lost"):
        assert st_m[k] == st_p[k], k
    assert st_m["lost"] == 0 and st_m["atoms"] > 2 * st_m["primaries_in"]
    scale = np.abs(out_p).max(axis=1, keepdims=True) + 1e-300
    assert (np.abs(out_m - out_p) / scale).max() < 1e-11


def test_mesh_slots_of_the_block_grid():
    """The reference's block grid: its Inner_sphere faces have the central cells behind
    them (followed, not crossed), the entry faces are the Extern, Input and Upper_wall
    faces, and the Axis faces are skipped."""
    from test_oracle_block import alpha, block
    from gelios_2d_amd.mesh import mesh_from_alpha
    m = mesh_from_alpha(alpha(block("b")))
    soff, sA, sB, code = ora.mc_mesh_slots(m)
    assert soff[-1] == len(code) == int(np.sum(m.nface))
    keep = (np.arange(len(m.fa) // m.ncells)[None, :] < m.nface[:, None]).ravel()
    t = m.ftype[keep]
    assert np.all(code[t == 4] == -3)
    assert np.all(code[t == 1] >= 0)
    assert np.all(code[(t == 2) | (t == 3) | (t == 5)] == -1)
    assert np.all(code[t == 0] >= 0) and not np.any(code == -2)


def test_mesh_walk_keeps_the_inflow_on_the_block_grid():
    """Liouville on the reference's own grid, in vacuum: every primary leaves, none is
    lost, and the density, velocity and temperature of the inflow hold in the volume
    mean and in the outer shells (2e5 primaries)."""
    from test_oracle_block import alpha, block
    from gelios_2d_amd.mesh import mesh_from_alpha
    m = mesh_from_alpha(alpha(block("b")))
    prim = np.zeros((5, m.ncells))
    prim[0] = prim[1] = 1e-200
    out, st = ora.mc_run_mesh(m, prim, ora.mc_population(m.cell_type, m.ncells), mc_reference_params(200_000, seed=3))
    assert st["lost"] == 0 and st["atoms"] == st["primaries_in"] > 150_000
    vol = np.zeros(m.ncells)
    for c in range(m.ncells):
        q = m.corners[m.coff[c]:m.coff[c + 1]]
        px, py = m.x[q], m.y[q]
        qx, qy = np.roll(px, -1), np.roll(py, -1)
        vol[c] = abs(((py + qy) * (px * qy - py * qx)).sum()) / 6.0 * 2 * np.pi
    cx, cy = m.cell_centers()
    outer = np.hypot(cx, cy) > 2.0
    n, u, T = out[25], out[26], out[28]
    assert abs(np.average(n, weights=vol) - 1.0) < 0.01
    assert abs(np.average(n[outer], weights=vol[outer]) - 1.0) < 0.01
    assert abs(np.average(u[outer], weights=vol[outer]) - V_INF) < 0.02
    assert abs(np.average(T[outer], weights=vol[outer]) - 1.0) < 0.02
