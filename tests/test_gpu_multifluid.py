"""The plasma + four hydrogen fluids sweep on the device (GPU tests): against the compiled
reference's fixtures (tests/golden/sweep_multifluid.npz) and the CPU oracle.

Tolerance: 1e-12 relative per conserved variable of every fluid (ro, ro u, ro v, E and
the plasma tracer), max |a - b| / max |b| after N steps, as for the single-fluid sweep;
the device's log and pow may differ from glibc's in the last bit.
"""
import numpy as np
import pytest

from conftest import golden
from gelios_2d_amd.params import GEL_FLUX_HLL, GEL_FLUX_HLLC

pytestmark = pytest.mark.gpu

TOL = 1e-12
GAM = 5.0 / 3.0


def conserved(prim):
    out = []
    for o in (0, 5, 9, 13, 17):
        ro, p, u, v = prim[o], prim[o + 1], prim[o + 2], prim[o + 3]
        out += [ro, ro * u, ro * v, p / (GAM - 1) + 0.5 * ro * (u * u + v * v)]
        if o == 0:
            out.append(prim[4])
    return out


def rel_err(a, b):
    return max(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300) for x, y in zip(conserved(a), conserved(b)))


def _fixture():
    from gelios_2d_amd.grid import PolarGrid
    from gelios_2d_amd.params import GEL_PHYS_GAS4H, reference_params
    d = golden("sweep_multifluid.npz")
    g = PolarGrid(int(d["ni"]), int(d["nj"]), d["x"], d["y"], d["cell_type"], d["outer_type"])
    prm = reference_params()
    prm.physics = GEL_PHYS_GAS4H
    prm.flux = GEL_FLUX_HLLC
    return d, g, prm


def test_static_matches_reference_fixture():
    from gelios_2d_amd.setka import Setka
    d, g, prm = _fixture()
    print("multifluid static fixture", flush=True)
    s = Setka(g, prm)
    s.upload(d["static_prim0"])
    st = s.Go_5_komponent__MK2(int(d["static_steps"]))
    assert st.n_numeric == 0
    assert rel_err(s.download(), d["static_prim"]) <= TOL


def test_moving_matches_reference_fixture():
    from gelios_2d_amd.setka import Setka
    d, g, prm = _fixture()
    print("multifluid moving fixture", flush=True)
    s = Setka(g, prm)
    s.upload(d["moving_prim0"])
    for k in range(int(d["moving_steps"])):
        s.set_dt(prm.dt0)       # the fixture calls Go_5_komponent_2(1) per step
        x2, y2 = d["traj"][k]
        s.step_moving(x2, y2)
    assert rel_err(s.download(), d["moving_prim"]) <= TOL


@pytest.mark.parametrize("flux", [GEL_FLUX_HLL, GEL_FLUX_HLLC])
def test_matches_oracle_ragged(oracle_lib, flux):
    """40 x 70 cells (two wavefront strips, the second ragged), rows in chunks of 7,
    every hydrogen region, 20 steps."""
    from gelios_2d_amd.configs import multifluid_heliosphere
    from gelios_2d_amd.setka import Setka
    w = multifluid_heliosphere(40, 70, 20, seed=3, flux=flux)
    print(f"multifluid ragged, flux {flux}", flush=True)
    ref = oracle_lib.OracleSweep.from_polar(w.grid, w.params).run(w.prim0, w.steps)
    s = Setka(w.grid, w.params)
    s.set_rows_per_wave(7)
    s.upload(w.prim0)
    st = s.Go_5_komponent__MK2(w.steps)
    assert st.n_numeric == 0
    assert rel_err(s.download(), ref) <= TOL


def test_reference_checkpoint_drives_the_sweep(oracle_lib, tmp_path):
    """A checkpoint written by the reference (Save_Setka_ALL_ALPHA) read into the device
    sweep: the plasma + hydrogen state and the grid it holds, 10 steps, against the
    oracle from the same parsed data."""
    from gelios_2d_amd import checkpoint as ck
    from gelios_2d_amd.params import GEL_PHYS_GAS4H, reference_params
    from gelios_2d_amd.setka import Setka
    d = golden("checkpoint_polar.npz")
    p = tmp_path / "ref.txt"
    p.write_bytes(d["ref_text"].tobytes())
    g, prim, _ = ck.polar_from_alpha(ck.read_alpha(p))
    prm = reference_params()
    prm.physics = GEL_PHYS_GAS4H
    prm.flux = GEL_FLUX_HLLC
    print("checkpoint-driven multifluid", flush=True)
    ref = oracle_lib.OracleSweep.from_polar(g, prm).run(prim, 10)
    s = Setka(g, prm)
    s.upload(prim)
    s.Go_5_komponent__MK2(10)
    assert rel_err(s.download(), ref) <= TOL
