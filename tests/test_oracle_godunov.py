"""The exact Riemann solver of the oracle (Solvers::Godunov_Solver_Alexashov,
Solvers.cpp:693-1321): bit for bit against vectors from the compiled reference
(gamma 5/3, all four wave patterns, the low-pressure-side swap, moving faces), and
against closed forms: Toro's exact star states for his five test problems (gamma 1.4),
consistency for equal states, mirror antisymmetry."""
import numpy as np
import pytest

from conftest import golden

GAM = 5.0 / 3.0


def _patterns(inp, g=GAM):
    g1, g2, gm = g - 1, g + 1, ((g - 1) / g) / 2
    out = []
    for r in inp:
        n = r[10:13]
        enI, enII = n @ r[2:5], n @ r[7:10]
        pI, pII, rI, rII = r[1], r[6], r[0], r[5]
        if pI > pII:
            enI, enII, pI, pII, rI, rII = -enII, -enI, pII, pI, rII, rI
        cI, cII = np.sqrt(g * pI / rI), np.sqrt(g * pII / rII)
        a = np.sqrt(rI * (g2 * pII + g1 * pI) / 2)
        Uud, Urz, Uvk = (pII - pI) / a, -2 * cII / g1 * (1 - (pI / pII) ** gm), -2 * (cII + cI) / g1
        Udf = enI - enII
        out.append(0 if Udf < Uvk else 1 if Udf <= Urz else 2 if Udf <= Uud else 3)
    return np.array(out)


def test_matches_reference_vectors(oracle_lib):
    d = golden("godunov_vectors.npz")
    inp, out = d["inp"], d["out"]
    counts = np.bincount(_patterns(inp), minlength=4)
    assert np.all(counts >= 50), counts            # vacuum, two fans, shock + fan, two shocks
    assert (inp[:, 1] > inp[:, 6]).sum() > 200 and (inp[:, 13] != 0).sum() > 200
    for r, ref in zip(inp, out):
        q, (sl, sp, sc), t, st = oracle_lib.godunov(r[0:5], r[5:10], r[10:13], r[13], r[14], 0.1, GAM)
        assert st == 0
        assert np.array_equal(np.concatenate([q, [sl, sp, sc, t]]), ref)


# Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, Table 4.2 (gamma 1.4):
# (rhoL, uL, pL, rhoR, uR, pR) -> (p*, u*)
TORO = [
    ((1.0, 0.0, 1.0, 0.125, 0.0, 0.1), (0.30313, 0.92745)),
    ((1.0, -2.0, 0.4, 1.0, 2.0, 0.4), (0.00189, 0.0)),
    ((1.0, 0.0, 1000.0, 1.0, 0.0, 0.01), (460.894, 19.5975)),
    ((1.0, 0.0, 0.01, 1.0, 0.0, 100.0), (46.0950, -6.19633)),
    ((5.99924, 19.5975, 460.894, 5.99242, -6.19633, 46.0950), (1691.64, 8.68975)),
]


@pytest.mark.parametrize("case", range(len(TORO)))
def test_toro_star_states(oracle_lib, case):
    """u* is the contact speed dsc; a face moving with the contact sees no mass flux and
    a momentum flux equal to p*."""
    (rl, ul, pl, rr, ur, pr), (ps, us) = TORO[case]
    L, R = [rl, pl, ul, 0.0, 0.0], [rr, pr, ur, 0.0, 0.0]
    _, (_, _, sc), _, st = oracle_lib.godunov(L, R, [1.0, 0.0, 0.0], 0.0, 1.0, 0.1, 1.4)
    assert st == 0
    assert abs(sc - us) <= 1e-5 * max(1.0, abs(us))
    q, _, _, _ = oracle_lib.godunov(L, R, [1.0, 0.0, 0.0], sc, 1.0, 0.1, 1.4)
    assert abs(q[0]) <= 1e-12 * max(1.0, abs(q[1]))
    assert abs(q[1] - ps) <= 2e-5 * ps + 1e-5


def test_equal_states_give_the_physical_flux(oracle_lib):
    rng = np.random.default_rng(3)
    for _ in range(200):
        ro, p = rng.uniform(0.1, 3.0), rng.uniform(0.05, 3.0)
        v = rng.normal(0, 2.0, 3)
        a = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(a), np.sin(a), 0.0])
        w = rng.normal(0, 1.0)
        q, _, _, st = oracle_lib.godunov([ro, p, *v], [ro, p, *v], n, w, 1.0, 0.1, GAM)
        assert st == 0
        un = n @ v
        E = p / (GAM - 1) + 0.5 * ro * (v @ v)
        want = np.array([ro * (un - w), *(ro * (un - w) * v + p * n), (un - w) * E + un * p])
        assert np.allclose(q, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_mirror_antisymmetry(oracle_lib):
    """Exchanging the sides and reversing the normal and the face speed reverses the
    flux: F(L, R, n, w) = -F(R, L, -n, -w)."""
    d = golden("godunov_vectors.npz")
    for r in d["inp"][:400]:
        q1, _, _, s1 = oracle_lib.godunov(r[0:5], r[5:10], r[10:13], r[13], 1.0, 0.1, GAM)
        q2, _, _, s2 = oracle_lib.godunov(r[5:10], r[0:5], -r[10:13], -r[13], 1.0, 0.1, GAM)
        assert s1 == s2 == 0
        assert np.allclose(q1, -q2, rtol=1e-9, atol=1e-12 * max(1.0, np.abs(q1).max()))
