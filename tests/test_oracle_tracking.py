"""Fitted-surface tracking in the oracle: Setka::Move_surface (Setka.cpp:3378-3844), the
A rails of Setka::Move_Setka_Calculate_2 (Setka.cpp:4324-4381) and the sweep's surface
faces (exact solver, one-sided reconstruction, Setka.cpp:10287-10370; Gran.cpp:345-426),
bit for bit against the compiled reference on three tracked heliospheres
(tests/golden/sweep_tracked.npz: final state and node positions), plus properties the
geometry fixes: every node stays on its rail, the rails' inner blocks sit where the
spacing law puts them, the surfaces keep their radial order."""
import numpy as np
import pytest

from conftest import golden


def _case(d, tag):
    from gelios_2d_amd.configs import tracked_heliosphere, tracked_spherical_ts
    kr = tracked_spherical_ts if bool(d[f"spherical_ts_{tag}"]) else None
    c = tracked_heliosphere(int(d[f"nj_{tag}"]), tuple(int(v) for v in d[f"blocks_{tag}"]), int(d[f"steps_{tag}"]),
                            int(d[f"seed_{tag}"]), flux=0, outer_faces=int(d[f"outer_faces_{tag}"]), key_radii=kr)
    assert np.array_equal(c.prim0, d[f"prim0_{tag}"]) and np.array_equal(c.grid.x, d[f"x0_{tag}"])
    assert np.array_equal(c.rails, d[f"rails_{tag}"])
    return c


@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_tracked_sweep_matches_reference(oracle_lib, tag):
    d = golden("sweep_tracked.npz")
    c = _case(d, tag)
    o = oracle_lib.TrackedOracle(c.grid, c.params, c.tracking, c.rails)
    out = o.run(c.prim0, c.steps, nthreads=1)
    assert np.abs(d[f"x_{tag}"] - c.grid.x).max() > 1e-3          # the surfaces moved
    assert np.array_equal(out, d[f"prim_{tag}"])
    assert np.array_equal(o.x, d[f"x_{tag}"]) and np.array_equal(o.y, d[f"y_{tag}"])


def test_reference_averaging_differs_from_per_line_averaging(oracle_lib):
    """Case c (spherical inner shock) moves both nodes of the inner-shock faces, so the
    reference's averaging of the contact line in place of the shocks' own lines
    (Setka.cpp:3718-3731, :3829-3837) shows in the result."""
    d = golden("sweep_tracked.npz")
    c = _case(d, "c")
    t = c.tracking.copy()
    t.average_lines = 1
    o = oracle_lib.TrackedOracle(c.grid, c.params, t, c.rails)
    o.run(c.prim0, c.steps, nthreads=1)
    assert not np.array_equal(o.x, d["x_c"])


def test_nodes_stay_on_rails_and_surfaces_in_order(oracle_lib):
    d = golden("sweep_tracked.npz")
    c = _case(d, "b")
    g, t = c.grid, c.tracking
    o = oracle_lib.TrackedOracle(g, c.params, t, c.rails)
    o.run(c.prim0, 10, nthreads=1)
    x, y = o.x.reshape(g.ni + 1, g.nj + 1), o.y.reshape(g.ni + 1, g.nj + 1)
    r = np.hypot(x, y)
    ang = np.arctan2(y, x)
    assert np.abs(ang - c.rails[None, :]).max() < 1e-12
    assert np.all(np.diff(r, axis=0) > 0)                              # radii increase along every rail
    rows = t.rows
    assert np.all(r[rows[0]] < r[rows[1]]) and np.all(r[rows[1]] < r[rows[2]])
    # the inner block is fixed by the spacing law r_j = R1 + (j/n)^1.3 (R111 - R1)
    j = np.arange(t.inner_rail_nodes)
    want = t.r_inner + (j / t.inner_rail_nodes) ** t.p_inner * (t.r_mid - t.r_inner)
    assert np.allclose(r[: t.inner_rail_nodes], want[:, None], rtol=1e-13, atol=0)
    # the outer end sits on the outer sphere
    assert np.allclose(r[-1], t.r_outer, rtol=1e-14)


def test_no_surfaces_reduces_to_the_plain_sweep(oracle_lib):
    """With no node on a surface the sweep's surface branch never triggers: the tracked
    mesh with all node types cleared steps exactly as the plain moving sweep."""
    d = golden("sweep_tracked.npz")
    c = _case(d, "a")
    g = c.grid
    x2, y2, st = oracle_lib.track_nodes(c.tracking, g.ni, g.nj, g.x, g.y, c.prim0, c.rails, c.params.gamma,
                                        c.params.dt0)
    assert st == 0
    a = oracle_lib.OracleSweep(oracle_lib.pack_general(g), g.x, g.y, c.params, x2, y2)
    a.ptype = np.zeros(g.nnodes, np.int32)
    b = oracle_lib.OracleSweep(oracle_lib.pack_general(g), g.x, g.y, c.params, x2, y2)
    ra, _ = a.step(c.prim0, nthreads=1)
    rb, _ = b.step(c.prim0, nthreads=1)
    assert np.array_equal(ra, rb)
