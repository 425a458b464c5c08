"""The sweep's FP64 division and square root (den / qt / sqn in sweep_core.h, a shorter
form of the compiler's own IEEE sequences) against IEEE arithmetic on the host: bit for
bit over random operands of normal exponents, quotients and roots next to a rounding
midpoint, all-ones and power-of-two significands, and zero numerators."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(a, b):
    from gelios_2d_amd import _lib
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n = a.size
    q, q2, s = np.empty(n), np.empty(n), np.empty(n)
    _lib.check(_lib.lib().gel_fp64_div_sqrt(n, _lib.dptr(a), _lib.dptr(b), _lib.dptr(q), _lib.dptr(q2),
                                            _lib.dptr(s)))
    return q, q2, s


def _random(rng, n, emax):
    m = rng.uniform(1.0, 2.0, n)
    e = rng.integers(-emax, emax + 1, n)
    return np.ldexp(m, e)


def _near_midpoints(rng, n):
    """(a, b) with a / b within a few ulps of the midpoint between two doubles: a is b
    times the midpoint, rounded."""
    b = _random(rng, n, 200)
    q = _random(rng, n, 200)
    mid = q + 0.5 * (np.nextafter(q, np.inf) - q)
    a = b * mid
    return a, b


def _bits_equal(x, y):
    return np.array_equal(x.view(np.int64), y.view(np.int64))


def test_division_and_square_root_are_ieee():
    rng = np.random.default_rng(9)
    n = 1 << 21
    a = _random(rng, n, 300) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    b = _random(rng, n, 300) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    am, bm = _near_midpoints(rng, n)
    a = np.concatenate([a, am])
    b = np.concatenate([b, bm])
    # significands of all ones and powers of two, both operands
    ones = np.ldexp(2.0 - 2.0 ** -52, rng.integers(-100, 100, 4096))
    pw2 = np.ldexp(1.0, rng.integers(-100, 100, 4096))
    a = np.concatenate([a, ones, pw2, ones, np.zeros(64)])
    b = np.concatenate([b, pw2, ones, ones[::-1], _random(rng, 64, 50)])
    q, q2, s = _run(a, b)
    ref = a / b
    assert np.array_equal(q, ref)
    nz = a != 0
    assert _bits_equal(q[nz], ref[nz])
    ref2 = np.roll(a, -1) / b
    assert _bits_equal(q2[np.roll(nz, -1)], ref2[np.roll(nz, -1)])
    pos = a >= 0
    assert _bits_equal(s[pos], np.sqrt(a[pos]))


def test_square_root_next_to_midpoints():
    """Arguments whose roots lie next to a rounding midpoint, and exact squares."""
    rng = np.random.default_rng(10)
    r = _random(rng, 1 << 20, 300)
    mid = r + 0.5 * (np.nextafter(r, np.inf) - r)
    x = np.concatenate([mid * mid, r * r, np.nextafter(r * r, 0.0), np.nextafter(r * r, np.inf)])
    _, _, s = _run(x, np.ones_like(x))
    assert _bits_equal(s, np.sqrt(x))
