"""The NumPy restatement of the surrogate network: finite-difference pin of its
back-propagation, and the BF16 rounding helper against exact cases."""
import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_rounding_exact_cases():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], np.float32)
    r = mo.bf16(a)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3*2^-8 -> 1 + 2^-6
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0]


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mo.n_params())
    X = rng.normal(0, 1, (8, 14))
    T = rng.normal(0, 1, (8, 5))
    loss, g = mo.loss_and_grad(p, X, T, "exact")
    idx = rng.choice(p.size, 60, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grad(pp, X, T)[0] - mo.loss_and_grad(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])
    assert p.size == 52101
