"""Combinations the library does not run fail loudly with an error status (GelError),
never by running something else: no CPU path, no silent change of physics."""
import numpy as np
import pytest

from gelios_2d_amd._lib import GelError
from gelios_2d_amd.params import GEL_C_1, GEL_PREC_FP32_BULK, GEL_SRC_HYBRID, GEL_SRC_TABLE

pytestmark = pytest.mark.gpu


def _with_c1_rows(g):
    t = np.full(g.ncells, 1, np.int8)          # C_2
    t[: 2 * g.nj] = GEL_C_1
    g.cell_type = t
    return g


@pytest.mark.parametrize("kind", ["fp32-gas", "mhd"])
def test_wind_cells_need_fp64_gas(kind):
    """C_1 cells (the radial reconstruction and rotation of the wind) run in the FP64
    gas sweep only; other contexts with them are refused at creation."""
    from gelios_2d_amd.configs import heliosphere, mhd_heliosphere
    from gelios_2d_amd.setka import Setka
    w = heliosphere(12, 8, 1, seed=1) if kind == "fp32-gas" else mhd_heliosphere(12, 8, 1, seed=1, moving=False)
    if kind == "fp32-gas":
        w.params.precision = GEL_PREC_FP32_BULK
    with pytest.raises(GelError):
        Setka(_with_c1_rows(w.grid), w.params)


def test_plasma_hydrogen_is_fp64_only():
    from gelios_2d_amd.configs import multifluid_heliosphere
    from gelios_2d_amd.setka import Setka
    w = multifluid_heliosphere(12, 8, 1, seed=1)
    w.params.precision = GEL_PREC_FP32_BULK
    with pytest.raises(GelError):
        Setka(w.grid, w.params)


def test_inner_region_is_static():
    from gelios_2d_amd.configs import inner_region
    from gelios_2d_amd.setka import Setka
    w = inner_region(12, 8, 1, seed=1)
    s = Setka(w.grid, w.params)
    s.upload(w.prim0)
    with pytest.raises(GelError):
        s.set_motion(1e-3, 50, 0.0)
    s.Go_5_komponent__MK2(1)                   # the static step still runs
    s.close()


def test_step_needs_state_and_sources():
    from gelios_2d_amd.setka import Setka
    from gelios_2d_amd.surrogate import workload_with_hydrogen
    w, f = workload_with_hydrogen(16, 10, 1, 2.0, 0.1, 3, moving=False)
    s = Setka(w.grid, w.params)
    with pytest.raises(GelError):              # nothing uploaded
        s.Go_5_komponent__MK2(1)
    s.upload(w.prim0)
    with pytest.raises(GelError):              # GEL_SRC_MK1 without the hydrogen moments
        s.Go_5_komponent__MK2(1)
    s.set_sources(f)
    s.Go_5_komponent__MK2(1)
    s.close()


def test_refresh_judge_values():
    """judge 2 (sticky) needs the hybrid source mode; judge values outside 0..2 are refused."""
    from gelios_2d_amd.setka import Setka
    from gelios_2d_amd.surrogate import SourceModel, workload_with_hydrogen
    from gelios_2d_amd.train_surrogate import DEFAULT_OUT
    model = SourceModel.load(DEFAULT_OUT)
    w, f = workload_with_hydrogen(16, 10, 0, 2.0, 0.1, 3, moving=False)
    for mode, judge in ((GEL_SRC_TABLE, 2), (GEL_SRC_HYBRID, 3), (GEL_SRC_HYBRID, -1)):
        prm = w.params.copy()
        prm.source_mode = mode
        s = Setka(w.grid, prm)
        s.upload(w.prim0)
        s.set_sources(f)
        with pytest.raises(GelError):
            model.refresh(s, tol=1e-3, judge=judge)
        s.close()
    prm = w.params.copy()
    prm.source_mode = GEL_SRC_HYBRID
    s = Setka(w.grid, prm)
    s.upload(w.prim0)
    s.set_sources(f)
    assert 0 <= model.refresh(s, tol=1e-3, judge=2) <= s.ncells
    s.close()
