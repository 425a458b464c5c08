"""The inner-region solver on the device (GPU tests): against the compiled reference's
fixtures (tests/golden/sweep_inner.npz) and the CPU oracle, 1e-12 relative per
conserved variable as every FP64 sweep (device pow and log may differ from glibc's in
the last bit)."""
import numpy as np
import pytest

from conftest import golden
from gelios_2d_amd.params import GEL_FLUX_HLL, GEL_FLUX_HLLC

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_err(a, b, gamma=5.0 / 3.0):
    from gelios_2d_amd.state import rel_err_conserved
    return rel_err_conserved(a, b, gamma)


def _run(w, rows=0):
    from gelios_2d_amd.setka import Setka
    s = Setka(w.grid, w.params)
    if rows:
        s.set_rows_per_wave(rows)
    s.upload(w.prim0)
    if getattr(w, "src", None) is not None:
        s.set_sources(w.src)
    st = s.Go_5_komponent__MK2(w.steps)
    return s.download(), st


@pytest.mark.parametrize("tag", ["plain", "sources"])
def test_matches_reference_fixture(tag):
    from gelios_2d_amd.configs import inner_region
    d = golden("sweep_inner.npz")
    w = inner_region(int(d["ni"]), int(d["nj"]), int(d[f"{tag}_steps"]), int(d[f"{tag}_seed"]),
                     sources=(tag == "sources"))
    print(f"inner fixture {tag}", flush=True)
    out, st = _run(w)
    assert rel_err(out, d[f"{tag}_prim"]) <= TOL


@pytest.mark.parametrize("flux", [GEL_FLUX_HLL, GEL_FLUX_HLLC])
def test_matches_oracle_ragged(oracle_lib, flux):
    """90 x 75 cells (two wavefront strips, the second ragged), chunks of 11 rows."""
    from gelios_2d_amd.configs import inner_region
    w = inner_region(90, 75, 15, 7, flux=flux, sources=True)
    print(f"inner ragged, flux {flux}", flush=True)
    sw = oracle_lib.OracleSweep.from_polar(w.grid, w.params)
    sw.src = w.src
    ref = sw.run(w.prim0, w.steps)
    out, st = _run(w, rows=11)
    assert rel_err(out, ref) <= TOL
