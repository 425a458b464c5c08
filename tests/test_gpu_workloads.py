"""The batched ensemble and the learned-source 800 x 400 workload at their full sizes
and lengths: FP64 batched cases against the oracle, and the FP32 / learned forms
against FP64 with the analytic charge-exchange source every step."""
import numpy as np
import pytest

from gelios_2d_amd.params import GEL_PREC_FP64, GEL_SRC_HYBRID, GEL_SRC_MK1
from gelios_2d_amd.state import l1_rel, rel_err_conserved

pytestmark = pytest.mark.gpu


def _cases(specs, precision=None, source_mode=None):
    from gelios_2d_amd.setka import Setka
    out = []
    for w, f in specs:
        prm = w.params.copy()
        if precision is not None:
            prm.precision = precision
        if source_mode is not None:
            prm.source_mode = source_mode
        s = Setka(w.grid, prm)
        s.upload(w.prim0)
        s.set_sources(f)
        if prm.source_mode == GEL_SRC_HYBRID:
            s.set_source_table(np.zeros((5, s.ncells)))
        out.append(s)
    return out


def _model():
    from gelios_2d_amd.surrogate import SourceModel
    from gelios_2d_amd.train_surrogate import DEFAULT_OUT
    return SourceModel.load(DEFAULT_OUT)


def test_ensemble_fp64_batch_matches_oracle(oracle_lib):
    """All 64 ensemble cases batched in FP64 with the analytic source: the smallest and
    the largest case equal the oracle after 30 steps."""
    from gelios_2d_amd.batch import Batch
    from gelios_2d_amd.configs import ensemble_cases
    specs = ensemble_cases(64, seed=5, steps=30)
    cases = _cases(specs, GEL_PREC_FP64, GEL_SRC_MK1)
    b = Batch(cases)
    b.step(30, True)
    sizes = [w.grid.ncells for w, _ in specs]
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        w, f = specs[k]
        prm = w.params.copy()
        prm.precision = GEL_PREC_FP64
        prm.source_mode = GEL_SRC_MK1
        sw = oracle_lib.OracleSweep.from_polar(w.grid, prm)
        sw.src = np.ascontiguousarray(f, np.float64)
        ref = sw.run(w.prim0, 30)
        err = rel_err_conserved(cases[k].download(), ref)
        print(f"case {k} {w.grid.ni}x{w.grid.nj}: batched FP64 vs oracle {err:.3e}", flush=True)
        assert err <= 1e-12
    for s in cases:
        s.close()


def test_ensemble_learned_500_steps_against_fp64():
    """The metric's workload at its length: 64 cases, FP32 bulk with the learned source
    (hybrid, sticky judge, refreshed every 10 steps inside the graphs), 500 graph-replayed
    steps; density and pressure L1 within 1e-4 of FP64 with the analytic source every step."""
    from gelios_2d_amd.batch import Batch
    from gelios_2d_amd.configs import ensemble_cases
    specs = ensemble_cases(64, seed=5, steps=500)
    m = _model()
    a = _cases(specs)
    A = Batch(a)
    A.refresh(m, tol=1e-3, judge=True)
    A.set_source_model(m, 10, tol=1e-3, judge=2)
    A.step(500, True)
    nref, nfall = A.source_stats()
    assert nref == 50
    r = _cases(specs, GEL_PREC_FP64, GEL_SRC_MK1)
    R = Batch(r)
    R.step(500, True)
    num = np.zeros(2)
    den = np.zeros(2)
    for x, y in zip(a, r):
        X, Y = x.download(), y.download()
        assert np.isfinite(X).all()
        for k in (0, 1):
            num[k] += np.abs(X[k] - Y[k]).sum()
            den[k] += np.abs(Y[k]).sum()
    l1 = num / den
    print(f"ensemble 500 steps: L1 density {l1[0]:.3e} pressure {l1[1]:.3e}; "
          f"fallback fraction {nfall / (nref * sum(s.ncells for s in a)):.3f}", flush=True)
    assert l1[0] <= 1e-4 and l1[1] <= 1e-4
    for s in a + r:
        s.close()


def test_learned_800x400_1000_steps_against_fp64():
    """The learned-source form of the moving 800 x 400 HLLD workload over its 1000 steps:
    FP64 state, the network's sources refreshed every 10 steps (hybrid, sticky judge),
    against FP64 with the analytic source every step: density and pressure L1 within 1e-4;
    without any source the same run is far off."""
    from gelios_2d_amd.configs import learned_mhd_800x400
    from gelios_2d_amd.params import GEL_SRC_NONE
    from gelios_2d_amd.setka import Setka
    out = {}
    for mode in (GEL_SRC_HYBRID, GEL_SRC_MK1, GEL_SRC_NONE):
        w = learned_mhd_800x400(GEL_PREC_FP64, mode)
        s = Setka(w.grid, w.params)
        s.set_motion(*w.motion)
        s.upload(w.prim0)
        s.set_sources(w.src)
        if mode == GEL_SRC_HYBRID:
            m = _model()
            s.set_source_table(np.zeros((5, s.ncells)))
            m.refresh(s, tol=1e-3, judge=True)
            nfall = m.advance(s, w.steps, every=10, tol=1e-3)
            print(f"learned 800x400: flagged cells after the last refresh {nfall}/{s.ncells}", flush=True)
        else:
            s.Go_5_komponent__MK2(w.steps)
        st = s.sync()
        assert st.n_numeric == 0
        out[mode] = s.download()
        s.close()
    ref = out[GEL_SRC_MK1]
    l1 = [l1_rel(out[GEL_SRC_HYBRID], ref, k) for k in (0, 1)]
    l1_none = [l1_rel(out[GEL_SRC_NONE], ref, k) for k in (0, 1)]
    print(f"learned 800x400, 1000 steps: L1 density {l1[0]:.3e} pressure {l1[1]:.3e}; "
          f"no source {l1_none[0]:.3e} {l1_none[1]:.3e}", flush=True)
    assert l1[0] <= 1e-4 and l1[1] <= 1e-4
    assert l1_none[1] > 10 * l1[1]
