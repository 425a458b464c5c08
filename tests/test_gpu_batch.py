"""Batched persistent launch: identical to running each case on its own, with and
without HIP-graph replay, for every launch shape; autotuner restores the states."""
import numpy as np
import pytest

from gelios_2d_amd.configs import ensemble_cases
from gelios_2d_amd.params import GEL_PREC_FP64

pytestmark = pytest.mark.gpu


def _cases(n, precision=None):
    from gelios_2d_amd.setka import Setka
    out = []
    for w, f in ensemble_cases(n, seed=11, steps=0, precision=precision):
        s = Setka(w.grid, w.params)
        s.upload(w.prim0)
        s.set_sources(f)
        s.set_source_table(np.zeros((5, s.ncells)))
        out.append(s)
    return out


@pytest.mark.parametrize("precision", [None, GEL_PREC_FP64])
def test_batch_equals_single_cases(precision):
    from gelios_2d_amd.batch import Batch
    single = _cases(5, precision)
    for s in single:
        s.Go_5_komponent__MK2(14)
    ref = [s.download() for s in single]
    for rows, wg, graph in ((0, 0, True), (3, 37, False), (16, 4096, True)):
        cases = _cases(5, precision)
        b = Batch(cases)
        b.set_launch(rows, wg)
        b.step(8, graph)
        b.step(6, graph)
        for s, r in zip(cases, ref):
            assert np.array_equal(s.download(), r), (rows, wg, graph)
        b.close()


def test_graph_replays_match_eager_each_time():
    from gelios_2d_amd.batch import Batch
    a, b_ = _cases(3), _cases(3)
    A, B = Batch(a), Batch(b_)
    for k in range(4):
        A.step(6, True)
        B.step(6, False)
        for x, y in zip(a, b_):
            assert np.array_equal(x.download(), y.download()), k
        x0 = a[0].download()      # eager host work between replays


def test_batch_rejects_a_context_twice():
    from gelios_2d_amd._lib import GelError
    from gelios_2d_amd.batch import Batch
    cases = _cases(2)
    with pytest.raises(GelError):
        Batch([cases[0], cases[1], cases[0]])
    b = Batch(cases)
    b.step(2, False)
    b.close()


def test_autotune_restores_state():
    from gelios_2d_amd.batch import Batch
    cases = _cases(4)
    before = [s.download() for s in cases]
    b = Batch(cases)
    rows, wg, ms = b.autotune(2)
    assert rows > 0 and wg > 0 and ms > 0
    for s, x in zip(cases, before):
        assert np.array_equal(s.download(), x)


def _model():
    from gelios_2d_amd.surrogate import SourceModel
    from gelios_2d_amd.train_surrogate import DEFAULT_OUT
    return SourceModel.load(DEFAULT_OUT)


def test_batch_refresh_equals_per_case_refresh():
    from gelios_2d_amd.batch import Batch
    m = _model()
    a, b_ = _cases(4), _cases(4)
    na = sum(m.refresh(s, tol=1e-3, judge=True) for s in a)
    B = Batch(b_)
    nb = B.refresh(m, tol=1e-3, judge=True)
    assert na == nb
    for x, y in zip(a, b_):
        assert np.array_equal(x.get_source_table(), y.get_source_table())


@pytest.mark.parametrize("every", [3, 4])
@pytest.mark.parametrize("judge", [1, 2])
def test_refresh_inside_graph_equals_interleaved_calls(every, judge):
    """Sources refreshed by the batch every `every` steps (inside the captured graph)
    give the same states, tables and flags as stepping and refreshing call by call;
    judge 2 runs the refresh over the batch's list of cells still on the network."""
    from gelios_2d_amd.batch import Batch
    m = _model()
    a, b_ = _cases(3), _cases(3)
    A, B = Batch(a), Batch(b_)
    A.refresh(m)
    B.refresh(m)
    B.set_source_model(m, every, tol=1e-3, judge=judge)
    total = 2 * 12 + every + 1
    B.step(total, True)
    for k in range(total // every):
        A.step(every, False)
        A.refresh(m, judge=judge)
    A.step(total % every, False)
    for x, y in zip(a, b_):
        assert np.array_equal(x.download(), y.download())
        assert np.array_equal(x.get_source_table(), y.get_source_table())
        assert np.array_equal(x.get_source_flags(), y.get_source_flags())
    r, n = B.source_stats()
    assert r == total // every and 0 <= n



@pytest.mark.parametrize("call", [1, 7, 20])
def test_refresh_schedule_counts_the_batch_steps(call):
    """The refresh follows every `every`-th step the batch advances, however the steps
    are split into calls (one step at a time, `every` + 3 per call, or calls longer than
    a graph cycle): states, tables, flags and the refresh count equal stepping and
    refreshing by hand; calls of one length replay their graphs."""
    from gelios_2d_amd.batch import Batch
    every, total = 4, 62
    m = _model()
    a, b_ = _cases(3), _cases(3)
    A, B = Batch(a), Batch(b_)
    A.refresh(m)
    B.refresh(m)
    B.set_source_model(m, every, tol=1e-3, judge=2)
    done = 0
    while done < total:
        n = min(call, total - done)
        B.step(n, True)
        done += n
    for k in range(total // every):
        A.step(every, False)
        A.refresh(m, judge=2)
    A.step(total % every, False)
    for x, y in zip(a, b_):
        assert np.array_equal(x.download(), y.download())
        assert np.array_equal(x.get_source_table(), y.get_source_table())
        assert np.array_equal(x.get_source_flags(), y.get_source_flags())
    r, n = B.source_stats()
    assert r == total // every


def test_prepared_graph_runs_the_same_steps():
    """prepare(n) captures the graphs of a call of n steps without running them: the
    states are untouched until step(n), which then equals plain launches."""
    from gelios_2d_amd.batch import Batch
    a, b_ = _cases(3), _cases(3)
    A, B = Batch(a), Batch(b_)
    before = [s.download() for s in b_]
    B.prepare(20)
    for s, x in zip(b_, before):
        assert np.array_equal(s.download(), x)
    B.step(20, True)
    A.step(20, False)
    for x, y in zip(a, b_):
        assert np.array_equal(x.download(), y.download())


def test_batch_sticky_refresh_equals_per_case_refresh():
    """judge 2 over a batch (the list of (case, cell) still on the network, a persistent
    grid over its chunks) leaves every case the tables, flags and fallback count that
    the per-case refresh (tiles of 256 cells, each packed) gives."""
    from gelios_2d_amd.batch import Batch
    m = _model()
    a, b_ = _cases(4), _cases(4)
    A = Batch(a)
    A.refresh(m)
    for s in b_:
        m.refresh(s, tol=1e-3, judge=1)
    A.step(18, True)
    for s in b_:
        s.Go_5_komponent__MK2(18)
    na = A.refresh(m, judge=2)
    nb = sum(m.refresh(s, tol=1e-3, judge=2) for s in b_)
    assert na == nb
    for x, y in zip(a, b_):
        assert np.array_equal(x.download(), y.download())
        assert np.array_equal(x.get_source_flags(), y.get_source_flags())
        assert np.array_equal(x.get_source_table(), y.get_source_table())


def _moving_setkas(precision, source, period=None):
    """Three moving MHD heliospheres of different sizes, each breathing with its own
    amplitude and phase and one common period; with source, the analytic closure of
    four frozen hydrogen populations."""
    from gelios_2d_amd.configs import mhd_heliosphere
    from gelios_2d_amd.setka import Setka
    from gelios_2d_amd.surrogate import workload_with_hydrogen
    out = []
    for k, (ni, nj, seed) in enumerate(((60, 30, 3), (97, 45, 4), (130, 70, 5))):
        f = None
        if source:
            w, f = workload_with_hydrogen(ni, nj, 0, 2.0 + 0.3 * k, 0.1, seed, moving=True)
        else:
            w = mhd_heliosphere(ni, nj, 0, seed=seed)
        w.params.precision = precision
        amp, per, phase = w.motion
        s = Setka(w.grid, w.params)
        s.set_motion(amp * (1.0 + 0.5 * k), period or per, phase + 0.7 * k)
        s.upload(w.prim0)
        if f is not None:
            s.set_sources(f)
        out.append(s)
    return out


@pytest.mark.parametrize("precision,source", [(GEL_PREC_FP64, False), (GEL_PREC_FP64, True), (1, False)])
def test_moving_batch_equals_single_cases(precision, source):
    """A batch of moving cases (prescribed breathing, one period, own amplitudes and
    phases) is bit-identical to each case stepped alone, eagerly and by graph replay
    over more than one 150-step cycle (6 buffer / time-step phases x the 50-step
    motion period), and after the autotuner's trials."""
    from gelios_2d_amd.batch import Batch
    single = _moving_setkas(precision, source)
    for s in single:
        s.Go_5_komponent__MK2(40)
    ref40 = [s.download() for s in single]
    for s in single:
        s.Go_5_komponent__MK2(157)
    ref = [s.download() for s in single]
    cases = _moving_setkas(precision, source)
    b = Batch(cases)
    b.autotune(3)
    b.step(17, False)
    b.step(23, True)
    for s, r in zip(cases, ref40):
        assert np.array_equal(s.download(), r)
    b.step(157, True)
    for s, r in zip(cases, ref):
        assert np.array_equal(s.download(), r)
    b.close()


def test_batch_rejects_mixed_motion():
    from gelios_2d_amd._lib import GelError
    from gelios_2d_amd.batch import Batch
    moving = _moving_setkas(GEL_PREC_FP64, False)
    static = _cases(1, GEL_PREC_FP64)
    with pytest.raises(GelError):
        Batch([moving[0], static[0]])
    other = _moving_setkas(GEL_PREC_FP64, False, period=30)
    with pytest.raises(GelError):
        Batch([moving[0], other[1]])


def _typed_setkas(moving, source):
    """Gas heliospheres (HLLC, FP64) with C_1 wind rows and two C_centr cells, static or
    breathing; with source, the analytic closure of frozen hydrogen."""
    from gelios_2d_amd.configs import MOTION_AMP, MOTION_PERIOD, heliosphere, motion_phase
    from gelios_2d_amd.hydrogen import populations
    from gelios_2d_amd.params import GEL_C_1, GEL_C_CENTR, GEL_SRC_MK1
    from gelios_2d_amd.setka import Setka
    out = []
    for k, (ni, nj, seed) in enumerate(((50, 40, 21), (83, 67, 22), (120, 61, 23))):
        w = heliosphere(ni, nj, 0, seed=seed, flux=1)
        w.grid.cell_type[: (6 + 2 * k) * nj] = GEL_C_1
        w.grid.cell_type[[nj // 2, 3 * nj // 2]] = GEL_C_CENTR
        if source:
            w.params.source_mode = GEL_SRC_MK1
        s = Setka(w.grid, w.params)
        if moving:
            s.set_motion(MOTION_AMP * (1 + k), MOTION_PERIOD, motion_phase(seed))
        s.upload(w.prim0)
        if source:
            s.set_sources(populations(w.grid, n0=0.1, seed=seed))
        out.append(s)
    return out


@pytest.mark.parametrize("moving,source", [(False, False), (True, False), (False, True), (True, True)])
def test_typed_batch_equals_single_cases(moving, source):
    """Grids with C_1 (radial reconstruction, rotated wind) and C_centr (frozen) cells in
    one batch: bit-identical with each case stepped alone, eagerly and by graph."""
    from gelios_2d_amd.batch import Batch
    single = _typed_setkas(moving, source)
    for s in single:
        s.Go_5_komponent__MK2(61)
    ref = [s.download() for s in single]
    cases = _typed_setkas(moving, source)
    b = Batch(cases)
    b.step(19, False)
    b.step(42, True)
    for s, r in zip(cases, ref):
        assert np.array_equal(s.download(), r)
    b.close()


def _mf_setkas(moving):
    """Plasma + four hydrogen fluids (HLLC, FP64, C_1 .. C_5 source regions) on three
    grids of different sizes, static or breathing."""
    from gelios_2d_amd.configs import MOTION_AMP, MOTION_PERIOD, motion_phase, multifluid_heliosphere
    from gelios_2d_amd.setka import Setka
    out = []
    for k, (ni, nj, seed) in enumerate(((40, 30, 31), (71, 50, 32), (96, 63, 33))):
        w = multifluid_heliosphere(ni, nj, 0, seed=seed)
        s = Setka(w.grid, w.params)
        if moving:
            s.set_motion(MOTION_AMP * (1 + k), MOTION_PERIOD, motion_phase(seed))
        s.upload(w.prim0)
        out.append(s)
    return out


@pytest.mark.parametrize("moving", [False, True])
def test_multifluid_batch_equals_single_cases(moving):
    """Plasma + hydrogen cases in one batch (one source launch and five fluid sweeps per
    step over every case): bit-identical with each case stepped alone, eagerly and by
    graph replay."""
    from gelios_2d_amd.batch import Batch
    single = _mf_setkas(moving)
    for s in single:
        s.Go_5_komponent__MK2(37)
    ref = [s.download() for s in single]
    cases = _mf_setkas(moving)
    b = Batch(cases)
    b.step(11, False)
    b.step(26, True)
    for s, r in zip(cases, ref):
        assert np.array_equal(s.download(), r)
    b.close()


def _tracked_setkas(flux):
    """Three tracked heliospheres (fitted inner shock, contact and outer shock on A rails,
    gas, FP64) of different sizes."""
    from gelios_2d_amd.configs import tracked_heliosphere
    from gelios_2d_amd.setka import Setka
    out = []
    for k, (nj, blocks, seed) in enumerate(((16, (24, 5, 5, 6), 51), (97, (28, 12, 12, 20), 52),
                                             (61, (28, 12, 10, 14), 53))):
        c = tracked_heliosphere(nj, blocks, 0, seed, flux=flux)
        s = Setka(c.grid, c.params)
        s.upload(c.prim0)
        s.set_tracking(c.tracking, c.rails)
        out.append(s)
    return out


@pytest.mark.parametrize("flux", [0, 1])
def test_tracked_batch_equals_single_cases(flux):
    """Cases that track their fitted surfaces (batched surface-speed and rail launches,
    then the sweep with surface faces; the node buffers change roles every step): state
    and node positions bit-identical with each case stepped alone, eagerly, by graph
    replay and after the autotuner's trials."""
    from gelios_2d_amd.batch import Batch
    single = _tracked_setkas(flux)
    for s in single:
        s.Go_5_komponent__MK2(23)
    ref = [(s.download(), *s.get_nodes()) for s in single]
    cases = _tracked_setkas(flux)
    b = Batch(cases)
    b.autotune(2)
    b.step(5, False)
    b.step(18, True)
    for s, (r, x, y) in zip(cases, ref):
        gx, gy = s.get_nodes()
        assert np.array_equal(s.download(), r)
        assert np.array_equal(gx, x) and np.array_equal(gy, y)
    b.close()
