"""The hydrogen Monte Carlo on the device (gel_mc_run, csrc/mc.hip) against the compiled
reference's pieces (tests/golden/mc_vectors.npz), the CPU oracle (oracle/mc_oracle.c)
atom for atom, and the closed forms of the physics at sizes the oracle does not reach."""
import os

import numpy as np
import pytest

from gelios_2d_amd import mc
from gelios_2d_amd.configs import heliosphere
from gelios_2d_amd.grid import polar_grid
from gelios_2d_amd.params import GEL_SRC_MK1, mc_reference_params, reference_params  # noqa: F401
from gelios_2d_amd.setka import Setka
from oracle import oracle as ora

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "mc_vectors.npz")
V_INF = -2.54338


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_device_rates_match_reference_and_oracle(gold):
    ri, ref = gold["rates_in"], gold["rates"]
    r = mc.rates(ri[:, 0], ri[:, 1])
    # the device's exp and erf differ from the host's in the last bits; Velosity_2 cancels
    np.testing.assert_allclose(r[:, :3], ref[:, :3], rtol=1e-12, atol=1e-13)
    m = ri[:, 0] / ri[:, 1] <= 7.0
    rel = np.abs(r[m, 3:] - ref[m, 3:]) / np.maximum(np.abs(ref[m, 3:]), 1e-12)
    assert rel.max() < 2e-5
    np.testing.assert_allclose(r, ora.mc_rates(ri[:, 0], ri[:, 1]), rtol=1e-11, atol=1e-13)


def test_device_sampler_matches_reference(gold):
    out, state = mc.exchange(gold["xin"], gold["xseed"], gold["samples"].shape[1])
    np.testing.assert_allclose(out, gold["samples"], rtol=1e-12, atol=1e-13)
    assert np.array_equal(state, gold["state"])


def _p(n, seed):
    p = mc_reference_params(n, seed=seed)
    p.max_segments = 100_000      # an atom crosses a few hundred cells here
    return p


def _plasma(nc, kind):
    prim = np.zeros((5, nc))
    if kind == "vacuum":
        prim[0] = prim[1] = 1e-200
    elif kind == "equilibrium":
        prim[0] = prim[1] = 0.5
        prim[2] = V_INF
    else:
        prim[0], prim[1], prim[2], prim[3] = 0.3, 0.6, 0.5, 0.2
    return prim


def _ctx(g, prim):
    s = Setka(g, reference_params())
    s.upload(prim)
    return s


@pytest.mark.parametrize("kind", ["hot", "equilibrium"])
def test_device_tracer_matches_oracle(kind):
    """Same seeds, same sensors per atom: the device traces the oracle's atoms; the
    tallies differ only by the order of the atomic sums and the math library's last bits."""
    g = polar_grid(8, 6, 1.0, 10.0)
    prim = _plasma(g.ni * g.nj, kind)
    p = _p(20000, 5)
    out_d, st_d = mc.run_mc(_ctx(g, prim), _p(20000, 5))
    out_o, st_o = ora.mc_run(g, prim, ora.mc_population(None, g.ni * g.nj), p)
    print(kind, st_d, st_o)
    for k in ("primaries_in", "atoms", "segments", "lost"):
        assert st_d[k] == st_o[k], k
    assert st_d["lost"] == 0 and st_d["dropped"] == 0
    scale = np.abs(out_o).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(out_d - out_o) / scale
    assert err.max() < 1e-9, (np.argmax(err.max(axis=1)), err.max())


def _volumes(g):
    x, y = g.x.reshape(g.ni + 1, g.nj + 1), g.y.reshape(g.ni + 1, g.nj + 1)
    px = np.stack([x[:-1, :-1], x[1:, :-1], x[1:, 1:], x[:-1, 1:]])
    py = np.stack([y[:-1, :-1], y[1:, :-1], y[1:, 1:], y[:-1, 1:]])
    G = sum((py[k] + py[(k + 1) % 4]) * (px[k] * py[(k + 1) % 4] - py[k] * px[(k + 1) % 4]) for k in range(4))
    return (np.abs(G) / 6.0 * 2 * np.pi).ravel()


def _row_mean(g, f, vol):
    f, w = f.reshape(g.ni, g.nj), vol.reshape(g.ni, g.nj)
    return (f * w).sum(1) / w.sum(1)


def test_device_vacuum_large():
    """Liouville at 1.6e7 atoms on 40 x 20 cells: the inflow's density, velocity and
    temperature in every ring of cells."""
    g = polar_grid(40, 20, 1.0, 50.0)
    out, st = mc.run_mc(_ctx(g, _plasma(g.ni * g.nj, "vacuum")), _p(16_000_000, 3))
    print(st)
    assert st["lost"] == 0 and st["dropped"] == 0 and st["atoms"] == st["primaries_in"]
    vol = _volumes(g)
    n, u, v, T = out[25], out[26], out[27], out[28]
    assert abs(np.average(n, weights=vol) - 1.0) < 0.002
    assert np.abs(_row_mean(g, n, vol)[5:] - 1.0).max() < 0.01
    assert np.abs(_row_mean(g, u, vol)[5:] - V_INF).max() < 0.01
    assert np.abs(_row_mean(g, v, vol)[5:]).max() < 0.01
    assert np.abs(_row_mean(g, T, vol)[5:] - 1.0).max() < 0.015


def test_device_equilibrium_large():
    g = polar_grid(40, 20, 1.0, 50.0)
    out, st = mc.run_mc(_ctx(g, _plasma(g.ni * g.nj, "equilibrium")), _p(4_000_000, 4))
    print(st)
    assert st["lost"] == 0 and st["dropped"] == 0 and st["generations"] > 3
    vol = _volumes(g)
    n, u, T = out[25], out[26], out[28]
    assert abs(np.average(n, weights=vol) - 1.0) < 0.005
    assert np.abs(_row_mean(g, n, vol)[10:] - 1.0).max() < 0.02
    assert np.abs(_row_mean(g, u, vol)[10:] - V_INF).max() < 0.02
    assert np.abs(_row_mean(g, T, vol)[10:] - 1.0).max() < 0.03


def test_device_mc_feeds_the_sweep():
    """The Monte Carlo on a heliosphere's plasma installs its 19 source inputs in a
    GEL_SRC_MK1 context: stepping it equals stepping a twin given the same fields through
    gel_set_sources."""
    w = heliosphere(100, 50, 20, seed=2)
    prm = w.params
    prm.source_mode = GEL_SRC_MK1
    a = Setka(w.grid, prm)
    a.upload(w.prim0)
    a.set_sources(np.tile(np.array([1.0] * 16 + [1.0, 1.0, 1.0])[:, None], (1, a.ncells)))
    out, st = mc.run_mc(a, _p(200_000, 9))
    print(st)
    assert st["lost"] == 0 and st["dropped"] == 0
    assert np.all(np.isfinite(out))
    vol = _volumes(w.grid)
    assert vol[out[25] > 0.0].sum() > 0.999 * vol.sum()   # atoms reach all but the smallest cells
    k = out[16:19]
    assert np.all((k >= 0.33) & (k <= 3.0))
    b = Setka(w.grid, prm)
    b.upload(w.prim0)
    b.set_sources(out[:19])
    a.Go_5_komponent__MK2(w.steps)
    b.Go_5_komponent__MK2(w.steps)
    np.testing.assert_array_equal(a.download(), b.download())


# ---------------------------------------------------------------- general mesh ----

def _block_ctx(prim=None, source_mode=None, tag="b"):
    """A context on the reference's own block grid (the Setka(N1..M4) checkpoint), at rest."""
    from test_oracle_block import alpha, block
    from gelios_2d_amd.mesh import mesh_from_alpha
    d = block(tag)
    m = mesh_from_alpha(alpha(d))
    prm = reference_params()
    if source_mode is not None:
        prm.source_mode = source_mode
    s = Setka.from_mesh(m, prm)
    s.upload(d["prim0"] if prim is None else prim(m.ncells))
    return s, m


def _mesh_volumes(m):
    vol = np.zeros(m.ncells)
    for c in range(m.ncells):
        q = m.corners[m.coff[c]:m.coff[c + 1]]
        px, py = m.x[q], m.y[q]
        qx, qy = np.roll(px, -1), np.roll(py, -1)
        vol[c] = abs(((py + qy) * (px * qy - py * qx)).sum()) / 6.0 * 2 * np.pi
    return vol


@pytest.mark.parametrize("kind", ["hot", "equilibrium"])
def test_mesh_tracer_equals_polar_tracer(kind):
    """The polar grid given as a general mesh (its face slots in Cell::Grans order, nodes
    as each cell sees them): the mesh walk traces the polar walk's atoms, which the oracle
    pins atom for atom -- the same counts, the tallies equal up to the order of the atomic
    sums and the last bits of face crossings solved from the other end of the face."""
    from gelios_2d_amd.mesh import mesh_from_polar
    g = polar_grid(8, 6, 1.0, 10.0)
    prim = _plasma(g.ni * g.nj, kind)
    out_p, st_p = mc.run_mc(_ctx(g, prim), _p(20000, 5))
    s = Setka.from_mesh(mesh_from_polar(g), reference_params())
    s.upload(prim)
    out_m, st_m = mc.run_mc(s, _p(20000, 5))
    print(kind, st_m, st_p)
    for k in ("primaries_in", "atoms", "segments", "lost", "generations"):
        assert st_m[k] == st_p[k], k
    assert st_m["lost"] == 0 and st_m["dropped"] == 0
    scale = np.abs(out_p).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(out_m - out_p) / scale
    assert err.max() < 1e-9, (np.argmax(err.max(axis=1)), err.max())


@pytest.mark.parametrize("kind,tag", [("hot", "b"), ("equilibrium", "b"), ("hot", "a")])
def test_mesh_tracer_matches_oracle_on_the_block_grid(kind, tag):
    """The reference's own grids (both block-grid checkpoints): the device walks the
    oracle's atoms one for one through the cells (the central cells behind the inner
    sphere included) -- equal counts, tallies within 1e-9."""
    s, m = _block_ctx(lambda nc: _plasma(nc, kind), tag=tag)
    out_d, st_d = mc.run_mc(s, _p(20000, 5))
    out_o, st_o = ora.mc_run_mesh(m, _plasma(m.ncells, kind), ora.mc_population(m.cell_type, m.ncells), _p(20000, 5))
    print(kind, st_d, st_o)
    for k in ("primaries_in", "atoms", "segments", "lost"):
        assert st_d[k] == st_o[k], k
    assert st_d["lost"] == 0 and st_d["dropped"] == 0 and st_d["atoms"] > st_d["primaries_in"]
    scale = np.abs(out_o).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(out_d - out_o) / scale
    assert err.max() < 1e-9, (np.argmax(err.max(axis=1)), err.max())


def _shells(m, vol):
    cx, cy = m.cell_centers()
    r = np.hypot(cx, cy)
    edges = [0.0, 0.5, 1.0, 2.0, 3.0, 4.0, 5.0, 7.0]
    return [(r >= lo) & (r < hi) for lo, hi in zip(edges[:-1], edges[1:])]


@pytest.mark.parametrize("kind,n,tol", [("vacuum", 16_000_000, 0.01), ("equilibrium", 4_000_000, 0.02)])
def test_mesh_tracer_keeps_the_inflow_on_the_block_grid(kind, n, tol):
    """Liouville on the reference's own grid (600 cells: the central cells behind the inner
    sphere, the fitted-surface rows, the cylinder and back plane): in vacuum, and with
    exchange against a plasma that moves and is as hot as the inflow (secondaries drawn
    from the same Maxwellian), the hydrogen keeps the inflow's density, velocity and
    temperature in every shell of cells."""
    s, m = _block_ctx(lambda nc: _plasma(nc, kind))
    out, st = mc.run_mc(s, _p(n, 3))
    print(kind, st)
    assert st["lost"] == 0 and st["dropped"] == 0
    if kind == "vacuum":
        assert st["atoms"] == st["primaries_in"]
    else:
        assert st["generations"] > 3
    vol = _mesh_volumes(m)
    nn, u, v, T = out[25], out[26], out[27], out[28]
    assert abs(np.average(nn, weights=vol) - 1.0) < tol / 4
    for sh in _shells(m, vol):
        w = vol[sh]
        means = [np.average(f[sh], weights=w) for f in (nn, u, v, T)]
        print(f"shell {sh.sum():4d} cells, volume {w.sum():9.3f}: n {means[0]:.4f} u {means[1]:.4f} "
              f"v {means[2]:+.4f} T {means[3]:.4f}", flush=True)
        assert abs(means[0] - 1.0) < tol
        assert abs(means[1] - V_INF) < tol
        assert abs(means[2]) < tol
        assert abs(means[3] - 1.0) < 1.5 * tol


def test_mesh_mc_feeds_the_block_grid_sweep():
    """The Monte Carlo on the block grid's plasma installs its source inputs in a
    GEL_SRC_MK1 mesh context: stepping it equals stepping a twin given the same fields."""
    a, m = _block_ctx(source_mode=GEL_SRC_MK1)
    a.set_sources(np.ones((19, m.ncells)))
    out, st = mc.run_mc(a, _p(400_000, 9))
    print(st)
    assert st["lost"] == 0 and st["dropped"] == 0
    assert np.all(np.isfinite(out))
    assert np.all((out[16:19] >= 0.33) & (out[16:19] <= 3.0))
    b, _ = _block_ctx(source_mode=GEL_SRC_MK1)
    b.set_sources(out[:19])
    a.Go_5_komponent__MK2(20)
    b.Go_5_komponent__MK2(20)
    ra = a.download()
    assert np.all(np.isfinite(ra))
    np.testing.assert_array_equal(ra, b.download())
