"""Fitted-surface tracking on the device (GPU tests): surface node speeds, A-rail
redistribution and the sweep with exact-solver surface faces, against the compiled
reference's fixtures (tests/golden/sweep_tracked.npz) and the CPU oracle.

Tolerances: 1e-12 relative per conserved variable after N steps, as for every FP64
sweep, and 1e-12 of the largest coordinate for the node positions: the device's pow
(exact solver, radial rescaling) can differ from glibc's in the last bit, which moves
the exact solver's iteration by an ulp; the rest of the arithmetic rounds as the host.
"""
import numpy as np
import pytest

from conftest import golden
from gelios_2d_amd.params import GEL_FLUX_HLL, GEL_FLUX_HLLC

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_err(a, b, gamma=5.0 / 3.0):
    from gelios_2d_amd.state import rel_err_conserved
    return rel_err_conserved(a, b, gamma)


def node_err(x, y, xr, yr):
    return max(np.abs(x - xr).max(), np.abs(y - yr).max()) / max(np.abs(xr).max(), np.abs(yr).max())


def _fixture_case(d, tag):
    from gelios_2d_amd.configs import tracked_heliosphere, tracked_spherical_ts
    kr = tracked_spherical_ts if bool(d[f"spherical_ts_{tag}"]) else None
    c = tracked_heliosphere(int(d[f"nj_{tag}"]), tuple(int(v) for v in d[f"blocks_{tag}"]), int(d[f"steps_{tag}"]),
                            int(d[f"seed_{tag}"]), flux=0, outer_faces=int(d[f"outer_faces_{tag}"]), key_radii=kr)
    assert np.array_equal(c.prim0, d[f"prim0_{tag}"])
    return c


def _run(c, steps, rows=0):
    from gelios_2d_amd.setka import Setka
    s = Setka(c.grid, c.params)
    if rows:
        s.set_rows_per_wave(rows)
    s.upload(c.prim0)
    s.set_tracking(c.tracking, c.rails)
    st = s.Go_5_komponent__MK2(steps)
    x, y = s.get_nodes()
    return s.download(), x, y, st


@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_tracked_sweep_matches_reference_fixture(tag):
    d = golden("sweep_tracked.npz")
    c = _fixture_case(d, tag)
    print(f"tracked fixture {tag}: {c.grid.ni}x{c.grid.nj}, {c.steps} steps", flush=True)
    prim, x, y, st = _run(c, c.steps)
    assert st.n_numeric == 0
    assert rel_err(prim, d[f"prim_{tag}"]) <= TOL
    assert node_err(x, y, d[f"x_{tag}"], d[f"y_{tag}"]) <= TOL


@pytest.mark.parametrize("flux", [GEL_FLUX_HLL, GEL_FLUX_HLLC])
def test_tracked_sweep_matches_oracle_ragged(oracle_lib, flux):
    """97 node columns (two wavefront strips, the second ragged), 71 rows in chunks of
    7; HLLC exercises the nul_potok faces beside the surfaces."""
    from gelios_2d_amd.configs import tracked_heliosphere
    c = tracked_heliosphere(97, (28, 12, 12, 20), 15, 41, flux=flux)
    print(f"tracked ragged: {c.grid.ni}x{c.grid.nj}, flux {flux}", flush=True)
    o = oracle_lib.TrackedOracle(c.grid, c.params, c.tracking, c.rails)
    ref = o.run(c.prim0, c.steps)
    prim, x, y, st = _run(c, c.steps, rows=7)
    assert st.n_numeric == 0
    assert rel_err(prim, ref) <= TOL
    assert node_err(x, y, o.x, o.y) <= TOL


def test_tracking_off_restores_the_static_grid():
    """gel_set_tracking(NULL) stops the motion: the nodes stay where the last tracked
    step left them."""
    from gelios_2d_amd.configs import tracked_heliosphere
    from gelios_2d_amd.setka import Setka
    c = tracked_heliosphere(16, (24, 5, 5, 6), 5, 43)
    s = Setka(c.grid, c.params)
    s.upload(c.prim0)
    s.set_tracking(c.tracking, c.rails)
    s.Go_5_komponent__MK2(3)
    x1, y1 = s.get_nodes()
    s.set_tracking(None, None)
    s.Go_5_komponent__MK2(2)
    x2, y2 = s.get_nodes()
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)
    assert not np.array_equal(x1, c.grid.x)
