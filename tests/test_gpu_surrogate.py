"""The learned source on the GPU: device features, the source-table sweep mode, the
per-cell judge/fallback, and what the trained network gains (ablation)."""
import numpy as np
import pytest

from gelios_2d_amd.params import GEL_SRC_MK1, GEL_SRC_NONE, GEL_SRC_TABLE
from gelios_2d_amd.state import l1_rel, rel_err_conserved

pytestmark = pytest.mark.gpu


def _setka(w, fields, mode):
    from gelios_2d_amd.setka import Setka
    prm = w.params.copy()
    prm.source_mode = mode
    s = Setka(w.grid, prm)
    if w.motion:
        s.set_motion(*w.motion)
    s.upload(w.prim0)
    s.set_sources(fields)
    return s


def test_features_match_host():
    from gelios_2d_amd import surrogate as sg
    w, f = sg.workload_with_hydrogen(60, 40, 0, 2.0, 0.1, 3, moving=False)
    s = _setka(w, f, GEL_SRC_MK1)
    x = sg.features(s)
    ro, p, u, v = w.prim0[:4]
    cx, cy = w.grid.cell_centers()
    ref = [np.log(ro), u, v, np.log(p / ro)]
    for g in range(2):
        a, b = 2 * g, 2 * g + 1
        n = f[a] + f[b]
        ref += [np.log(n), (f[a] * f[4 + a] + f[b] * f[4 + b]) / n, (f[a] * f[8 + a] + f[b] * f[8 + b]) / n,
                np.log((f[a] * f[12 + a] + f[b] * f[12 + b]) / n)]
    ref += [cy, cx]
    ref = np.stack(ref, 1)
    assert np.allclose(x, ref, rtol=1e-5, atol=1e-5)


def test_table_mode_is_the_analytic_source_for_one_step():
    """A table holding the analytic sources of the current state reproduces one step
    of the sweep with the analytic closure (Cell::Get_Sourse_MK1, Cell.cpp:498-665)."""
    from gelios_2d_amd import surrogate as sg
    w, f = sg.workload_with_hydrogen(80, 64, 1, 2.5, 0.15, 5)
    a = _setka(w, f, GEL_SRC_MK1)
    b = _setka(w, f, GEL_SRC_TABLE)
    b.set_source_table(sg.analytic_sources(b))
    a.Go_5_komponent__MK2(1)
    b.Go_5_komponent__MK2(1)
    assert rel_err_conserved(b.download(), a.download()) <= 1e-14


def test_trained_network_judge_and_gain():
    from gelios_2d_amd import surrogate as sg
    X, Y = sg.training_data(ni=120, nj=60, n_sets=6, seed=4, max_evolve=120)
    model = sg.train_source_model(X, Y, steps=2400, batch=2048, lr=2e-3)
    assert model.loss_history[-100:].mean() < 0.05 * model.loss_history[:10].mean()

    # held-out inflow set
    w, f = sg.workload_with_hydrogen(120, 60, 60, 2.2, 0.12, 99)
    ana = _setka(w, f, GEL_SRC_MK1)
    net = _setka(w, f, GEL_SRC_TABLE)
    nfall = model.refresh(net, tol=1e-3, judge=True)
    assert 0 <= nfall < net.ncells

    # ablation over 60 steps, sources refreshed every 10 steps from the current state
    none = _setka(w, f, GEL_SRC_NONE)
    for k in range(6):
        ana.Go_5_komponent__MK2(10)
        model.refresh(net, tol=1e-3, judge=True)
        net.Go_5_komponent__MK2(10)
        none.Go_5_komponent__MK2(10)
    A, N, Z = ana.download(), net.download(), none.download()
    err_net = rel_err_conserved(N, A)
    err_none = rel_err_conserved(Z, A)
    print(f"fallback cells {nfall}/{net.ncells}; error vs analytic: network {err_net:.3e}, no source {err_none:.3e}")
    assert err_net < 0.1 * err_none


def _model():
    from gelios_2d_amd.surrogate import SourceModel
    from gelios_2d_amd.train_surrogate import DEFAULT_OUT
    return SourceModel.load(DEFAULT_OUT)


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_refresh_judge_at_zero_tolerance_is_the_analytic_table(precision):
    """With tol = 0 every cell falls back, so the fused refresh must leave exactly the
    analytic sources (Cell::Get_Sourse_MK1, Cell.cpp:590-664) of the current state."""
    from gelios_2d_amd import surrogate as sg
    w, f = sg.workload_with_hydrogen(70, 50, 0, 2.0, 0.1, 7, moving=False)
    w.params.precision = precision
    s = _setka(w, f, GEL_SRC_TABLE)
    n = _model().refresh(s, tol=0.0, judge=True)
    assert n == s.ncells
    assert np.array_equal(s.get_source_table(), sg.analytic_sources(s))


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_refresh_matches_unfused_forward(precision):
    """Network-only table (no judge) against the unfused MFMA forward (gel_mlp_forward)
    on the same normalised features: the layers, biases, SiLU and the output scaling
    agree; inputs that round to a different BF16 on the two paths (features normalised
    from FP64 on the device, from FP32 on the host) move single cells slightly."""
    from gelios_2d_amd import surrogate as sg
    w, f = sg.workload_with_hydrogen(90, 60, 0, 2.4, 0.12, 9, moving=False)
    w.params.precision = precision
    s = _setka(w, f, GEL_SRC_TABLE)
    m = _model()
    m.refresh(s, tol=1e-3, judge=False)
    got = s.get_source_table()
    x = ((sg.features(s).astype(np.float64) - m.mu) / m.sd).astype(np.float32)
    y = m.mlp.forward(x).astype(np.float64)
    prim = s.download()
    scale = w.params.src_coef * prim[0] * f[0:4].sum(axis=0)
    want = ((y * m.ysd.astype(np.float64) + m.ymu.astype(np.float64)) * scale[:, None]).T
    unit = np.maximum(m.ysd.astype(np.float64)[:, None] * scale[None, :], 1e-300)
    rel = np.abs(got - want) / unit
    assert np.median(rel) < 1e-6
    assert np.mean(rel < 1e-6) > 0.9
    assert rel.max() < 0.1


@pytest.mark.parametrize("precision", [0, 1])
def test_hybrid_mode_limits(precision):
    """GEL_SRC_HYBRID after a refresh with tol = 0 (every cell flagged) is the MK1 sweep
    bit for bit; with no judge (no cell flagged) it is the table sweep."""
    from gelios_2d_amd import surrogate as sg
    from gelios_2d_amd.params import GEL_SRC_HYBRID
    w, f = sg.workload_with_hydrogen(80, 64, 0, 2.2, 0.1, 13, moving=False)
    w.params.precision = precision
    m = _model()
    mk1 = _setka(w, f, GEL_SRC_MK1)
    hyb = _setka(w, f, GEL_SRC_HYBRID)
    hyb.set_source_table(np.zeros((5, hyb.ncells)))
    assert m.refresh(hyb, tol=0.0, judge=True) == hyb.ncells
    mk1.Go_5_komponent__MK2(12)
    hyb.Go_5_komponent__MK2(12)
    assert np.array_equal(mk1.download(), hyb.download())
    tab = _setka(w, f, GEL_SRC_TABLE)
    hyb2 = _setka(w, f, GEL_SRC_HYBRID)
    hyb2.set_source_table(np.zeros((5, hyb2.ncells)))
    m.refresh(tab, tol=1e-3, judge=False)
    assert m.refresh(hyb2, tol=1e-3, judge=False) == 0
    tab.Go_5_komponent__MK2(12)
    hyb2.Go_5_komponent__MK2(12)
    assert np.array_equal(tab.download(), hyb2.download())


def test_fp32_bulk_analytic_source_against_fp64():
    """The FP32-bulk static sweep with the analytic closure (moments in FP32, copied into
    LDS during the face solves, the closure in FP32) against the FP64 sweep with the
    reference's arithmetic, 120 steps on 90 x 70 cells (two strips, the second ragged):
    density and pressure L1 errors within 2e-5 (the FP32-bulk ensemble measures a few
    1e-6), and the source visibly acting (the run without it is 10x further off)."""
    from gelios_2d_amd import surrogate as sg
    w, f = sg.workload_with_hydrogen(90, 70, 0, 2.4, 0.15, 21, moving=False)
    runs = {}
    for name, prec, mode in (("fp64", 0, GEL_SRC_MK1), ("fp32", 1, GEL_SRC_MK1), ("none", 0, GEL_SRC_NONE)):
        w.params.precision = prec
        s = _setka(w, f, mode)
        s.Go_5_komponent__MK2(120)
        runs[name] = s.download()
    for k in (0, 1):
        err = l1_rel(runs["fp32"], runs["fp64"], k)
        gap = l1_rel(runs["none"], runs["fp64"], k)
        print(f"variable {k}: fp32 vs fp64 L1 {err:.3e}, no source vs fp64 {gap:.3e}")
        assert err <= 2e-5
        assert gap > 10 * err


@pytest.mark.parametrize("precision", [0, 1])
def test_sticky_judge_keeps_flags_and_matches_full_refresh(precision):
    """judge 2: a cell the previous refresh flagged keeps the analytic closure and is
    not evaluated; every other cell gets exactly what a full refresh (judge 1) of the
    same state gives it (table value and flag), though the tile packs it elsewhere."""
    from gelios_2d_amd import surrogate as sg
    from gelios_2d_amd.params import GEL_SRC_HYBRID
    w, f = sg.workload_with_hydrogen(90, 70, 0, 2.3, 0.12, 17, moving=False)
    w.params.precision = precision
    m = _model()
    a = _setka(w, f, GEL_SRC_HYBRID)
    a.set_source_table(np.zeros((5, a.ncells)))
    m.refresh(a, tol=1e-3, judge=1)
    flags0 = a.get_source_flags()
    assert 0 < flags0.sum() < a.ncells
    a.Go_5_komponent__MK2(25)
    n2 = m.refresh(a, tol=1e-3, judge=2)
    ta, fa = a.get_source_table(), a.get_source_flags()
    w.prim0 = a.download()
    b = _setka(w, f, GEL_SRC_HYBRID)
    b.set_source_table(np.zeros((5, b.ncells)))
    m.refresh(b, tol=1e-3, judge=1)
    tb, fb = b.get_source_table(), b.get_source_flags()
    live = flags0 == 0
    assert np.array_equal(fa, np.where(live, fb, 1))
    assert n2 == int(fa.sum())
    assert np.array_equal(ta[:, live], tb[:, live])
