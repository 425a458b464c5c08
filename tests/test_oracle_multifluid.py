"""The plasma + four hydrogen fluids sweep of the oracle (Setka::Go_5_komponent_2,
Setka.cpp:8951-9777) bit for bit against the compiled reference on the 16 x 10
harness grid with every source region (tests/golden/sweep_multifluid.npz: 30 static
steps, 6 moving steps), and a reduction the equations fix: without charge exchange the
plasma of the multi-fluid step is the single-fluid step's plasma."""
import numpy as np

from conftest import golden


def _setup(d):
    from gelios_2d_amd.grid import PolarGrid, pack_general
    from gelios_2d_amd.params import GEL_FLUX_HLLC, GEL_PHYS_GAS4H, reference_params
    g = PolarGrid(int(d["ni"]), int(d["nj"]), d["x"], d["y"], d["cell_type"], d["outer_type"])
    prm = reference_params()
    prm.physics = GEL_PHYS_GAS4H
    prm.flux = GEL_FLUX_HLLC                       # met = 1, Setka.cpp:9011
    return g, prm, pack_general(g)


def test_static_matches_reference(oracle_lib):
    d = golden("sweep_multifluid.npz")
    g, prm, mesh = _setup(d)
    sw = oracle_lib.OracleSweep(mesh, g.x, g.y, prm)
    out = sw.run(d["static_prim0"], int(d["static_steps"]), nthreads=1)
    assert not np.array_equal(out, d["static_prim0"])
    assert np.array_equal(out, d["static_prim"])


def test_moving_matches_reference(oracle_lib):
    d = golden("sweep_multifluid.npz")
    g, prm, mesh = _setup(d)
    sw = oracle_lib.OracleSweep(mesh, g.x, g.y, prm)
    cur, x, y = d["moving_prim0"].copy(), g.x, g.y
    for s in range(int(d["moving_steps"])):
        x2, y2 = d["traj"][s]
        sw.set_nodes(x, y, x2, y2)
        sw.dt = prm.dt0              # the fixture calls Go_5_komponent_2(1) per step
        cur, st = sw.step(cur, nthreads=1)
        assert st == 0
        x, y = x2, y2
    assert np.array_equal(cur, d["moving_prim"])


def test_no_exchange_plasma_is_the_single_fluid_step(oracle_lib):
    d = golden("sweep_multifluid.npz")
    g, prm, mesh = _setup(d)
    prm.src_coef = 0.0
    prm.src_coef_h = 0.0
    mf = oracle_lib.OracleSweep(mesh, g.x, g.y, prm)
    p5 = prm.copy()
    p5.physics = 0
    sf = oracle_lib.OracleSweep(mesh, g.x, g.y, p5)
    a, _ = mf.step(d["static_prim0"], nthreads=1, dt_prev=1e-3)
    b, _ = sf.step(d["static_prim0"][:5], nthreads=1, dt_prev=1e-3)
    assert np.array_equal(a[:5], b)
    assert not np.array_equal(a[5:], d["static_prim0"][5:])
