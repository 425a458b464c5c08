"""The no-allocation flux entry point (gel_flux_hlld_batch_noalloc) against the
allocating one (gel_flux_hlld_batch) through the C-ABI: identical fluxes and wave
speeds on two calls in a row, the second with fewer faces, and sizes beyond the
staging array's limit rejected."""
import ctypes

import numpy as np
import pytest

from gelios_2d_amd import _lib

pytestmark = pytest.mark.gpu

MAX_FACES = 32768   # GEL_NOALLOC_MAX_FACES


def _faces(n, seed):
    rng = np.random.default_rng(seed)
    L = np.stack([rng.uniform(0.05, 3, n), rng.uniform(0.01, 3, n), *rng.normal(0, 1.5, (3, n)),
                  *rng.normal(0, 2.0, (3, n)), rng.uniform(0, 2, n)], 1)
    R = np.stack([rng.uniform(0.05, 3, n), rng.uniform(0.01, 3, n), *rng.normal(0, 1.5, (3, n)),
                  *rng.normal(0, 2.0, (3, n)), rng.uniform(0, 2, n)], 1)
    a = rng.uniform(0, 2 * np.pi, n)
    W = np.where(rng.random(n) < 0.5, 0.0, rng.normal(0, 1.0, n))
    return [np.ascontiguousarray(v, np.float64) for v in (L, R, W, np.cos(a), np.sin(a))]


def _call(name, n, L, R, W, n1, n2):
    P = np.full((n, 9), np.nan)
    smax = np.full(n, np.nan)
    st = getattr(_lib.lib(), name)(n, *[_lib.dptr(v) for v in (L, R, W, n1, n2)], 5.0 / 3.0, _lib.dptr(P),
                                   _lib.dptr(smax))
    return st, P, smax


def test_noalloc_flux_equals_allocating_flux():
    for n, seed in ((MAX_FACES, 1), (1000, 2)):
        args = _faces(n, seed)
        st_a, P_a, s_a = _call("gel_flux_hlld_batch", n, *args)
        st_b, P_b, s_b = _call("gel_flux_hlld_batch_noalloc", n, *args)
        assert st_a == 0 and st_b == 0
        assert np.array_equal(P_a, P_b)
        assert np.array_equal(s_a, s_b)


def test_noalloc_flux_rejects_more_faces_than_its_array():
    n = MAX_FACES + 1
    args = _faces(n, 3)
    st, P, _ = _call("gel_flux_hlld_batch_noalloc", n, *args)
    assert st == 1                                   # GEL_E_ARG
    assert b"GEL_NOALLOC_MAX_FACES" in _lib.lib().gel_last_error()
    assert np.isnan(P).all()
