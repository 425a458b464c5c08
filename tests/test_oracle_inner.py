"""The inner-region solver of the oracle (Setka::Go_stationary_5_komponent_inner_MK2,
Setka.cpp:6999-7212) bit for bit against the compiled reference on a 24 x 12 inner
region (tests/golden/sweep_inner.npz: 40 steps, and 25 steps with the MK1 closure of
seeded hydrogen moments), and properties of its update: the tracer is passive and
loses a bare dt per step (Setka.cpp:7173)."""
import numpy as np
import pytest

from conftest import golden


@pytest.mark.parametrize("tag", ["plain", "sources"])
def test_matches_reference(oracle_lib, tag):
    from gelios_2d_amd.configs import inner_region
    d = golden("sweep_inner.npz")
    w = inner_region(int(d["ni"]), int(d["nj"]), int(d[f"{tag}_steps"]), int(d[f"{tag}_seed"]),
                     sources=(tag == "sources"))
    assert np.array_equal(w.prim0, d[f"{tag}_prim0"])
    sw = oracle_lib.OracleSweep.from_polar(w.grid, w.params)
    if tag == "sources":
        assert np.array_equal(w.src, d[f"{tag}_src"])
        sw.src = w.src
    out = sw.run(w.prim0, w.steps, nthreads=1)
    assert not np.array_equal(out, w.prim0)
    assert np.array_equal(out, d[f"{tag}_prim"])


def test_tracer_is_passive_and_loses_a_bare_dt(oracle_lib):
    """With no tracer anywhere (cells, and the inner sphere's ghost) the tracer flux is
    zero, so one step leaves Q = -dt exactly in every updated cell (the bare dt of
    Setka.cpp:7173) and 0 in the skipped C_centr cells; and the tracer is passive: the
    other fields do not depend on it."""
    from gelios_2d_amd.configs import inner_region
    from gelios_2d_amd.params import GEL_C_CENTR
    w = inner_region(16, 10, 1, 5)
    w.params.in_q = 0.0
    a = w.prim0.copy()
    a[4] = 0.0
    dt = 1e-4
    sw = oracle_lib.OracleSweep.from_polar(w.grid, w.params)
    out_a, _ = sw.step(a, nthreads=1, dt_prev=dt)
    frozen = w.grid.cell_type == GEL_C_CENTR
    assert frozen.any()
    assert np.all(out_a[4][~frozen] == -dt) and np.all(out_a[4][frozen] == 0.0)
    out_b, _ = sw.step(w.prim0, nthreads=1, dt_prev=dt)
    assert np.array_equal(out_a[:4], out_b[:4])
