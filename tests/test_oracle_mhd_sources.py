"""The geometric source terms of the axisymmetric MHD update (swirl w, azimuthal field
B_phi), which have no counterpart in the reference (HLLDQ_Korolkov, Solvers.cpp:1642-2185,
is a flux only).  They are pinned here by what the mathematics fixes:

* the residual of one step on smooth fields converges, as the grid is refined, to the
  right-hand side of the axisymmetric ideal-MHD equations in r-weighted conservation
  form, d(r U)/dt + d(r F_z)/dz + d(r F_r)/dr = S, written out independently below
  (fluxes from the textbook forms, derivatives by central differences of the exact
  flux functions);
* rigidly rotating gas (rho w^2 / r = dp/dr) and a B_phi pinch
  (d(p + B_phi^2 / 8 pi)/dr = -B_phi^2 / (4 pi r)) are steady to truncation order;
* flipping (w, B_phi) -> (-w, -B_phi) flips them in the solution and changes nothing else.

A wrong sign, a dropped term or a wrong factor in any of the three sources leaves an
O(1) residual in its variable that does not shrink with the grid (checked when these
tests were written by building the oracle with each such mistake).
"""
import numpy as np
import pytest

from gelios_2d_amd.grid import polar_grid
from gelios_2d_amd.params import GEL_FLUX_HLLC, GEL_PHYS_MHD, heliosphere_params

G = 5.0 / 3.0
PI4 = 4.0 * np.pi


def cons(q):
    ro, p, u, v, w, bx, by, bz, Q = q
    B2 = bx * bx + by * by + bz * bz
    E = p / (G - 1) + 0.5 * ro * (u * u + v * v + w * w) + B2 / (2 * PI4)
    return np.array([ro, ro * u, ro * v, ro * w, E, bx, by, bz, Q])


def flux(q, nx, ny):
    """Ideal-MHD flux along (nx, ny) in the meridional plane, Gaussian B."""
    ro, p, u, v, w, bx, by, bz, Q = q
    B2 = bx * bx + by * by + bz * bz
    pT = p + B2 / (2 * PI4)
    E = p / (G - 1) + 0.5 * ro * (u * u + v * v + w * w) + B2 / (2 * PI4)
    un, bn, ub = u * nx + v * ny, bx * nx + by * ny, u * bx + v * by + w * bz
    return np.array([ro * un, ro * un * u + pT * nx - bn * bx / PI4, ro * un * v + pT * ny - bn * by / PI4,
                     ro * un * w - bn * bz / PI4, (E + pT) * un - bn * ub / PI4, un * bx - u * bn,
                     un * by - v * bn, un * bz - w * bn, Q * un])


def geometric_source(q):
    """S of the r-weighted form: r-momentum rho w^2 + p_T - B_phi^2/4pi, phi-momentum
    -(rho u_r w - B_r B_phi/4pi), B_phi u_r B_phi - w B_r."""
    ro, p, u, v, w, bx, by, bz, Q = q
    pT = p + (bx * bx + by * by + bz * bz) / (2 * PI4)
    z = np.zeros_like(ro)
    return np.array([z, z, ro * w * w + pT - bz * bz / PI4, -(ro * v * w - by * bz / PI4), z, z, z,
                     v * bz - w * by, z])


def rhs(fields, x, y, h=1e-5):
    """dU/dt = (S - d(y F_x)/dx - d(y F_y)/dy) / y at (x, y)."""
    fx = lambda a, b: b * flux(fields(a, b), 1.0, 0.0)
    fy = lambda a, b: b * flux(fields(a, b), 0.0, 1.0)
    div = (fx(x + h, y) - fx(x - h, y)) / (2 * h) + (fy(x, y + h) - fy(x, y - h)) / (2 * h)
    return (geometric_source(fields(x, y)) - div) / y


def smooth(x, y):
    """Smooth fields with swirl, a divergence-free poloidal field and B_phi."""
    return np.array([1 + 0.2 * np.sin(0.7 * x) * np.cos(0.5 * y), 1 + 0.1 * np.cos(0.6 * x + 0.3 * y),
                     0.3 * np.sin(0.4 * y) + 0.1 * np.cos(0.5 * x), 0.2 * np.cos(0.3 * x) * np.sin(0.2 * y),
                     0.5 + 0.3 * np.sin(0.5 * x) * np.cos(0.3 * y), 0.8 + 0.05 * np.sin(0.5 * x) / y,
                     -0.025 * np.cos(0.5 * x), 0.4 + 0.2 * np.cos(0.3 * x) * np.sin(0.4 * y),
                     1 + 0.1 * np.sin(x)])


def rigid_rotation(x, y, om=0.6):
    """rho = 1, w = om r, p = 1 + om^2 r^2 / 2: the pressure gradient holds the swirl."""
    z = np.zeros_like(x)
    return np.array([1 + z, 1 + 0.5 * om * om * y * y, z, z, om * y, z, z, z, 1 + z])


def pinch(x, y, b0=1.2):
    """B_phi = b0 r, p = 4 - b0^2 r^2 / (4 pi): magnetic pressure and tension balance the
    pressure gradient."""
    z = np.zeros_like(x)
    return np.array([1 + z, 4 - b0 * b0 * y * y / PI4, z, z, z, z, z, b0 * y, 1 + z])


def residual(oracle_lib, fields, ni, nj, dt=1e-7):
    """One-step residual (U' - U)/dt of the oracle on a static polar grid r in [2, 4] and
    the exact right-hand side, on the cells at least 3 cells from the boundaries and
    with r sin(theta) > 1.2."""
    g = polar_grid(ni, nj, 2.0, 4.0, spacing="uniform")
    prm = heliosphere_params(flux=GEL_FLUX_HLLC)
    prm.physics = GEL_PHYS_MHD
    cx, cy = g.cell_centers()
    q0 = fields(cx, cy)
    q1, st = oracle_lib.OracleSweep.from_polar(g, prm).step(q0, dt_prev=dt)
    assert st == 0
    I, J = np.meshgrid(np.arange(ni), np.arange(nj), indexing="ij")
    m = ((I >= 3) & (I < ni - 3) & (J >= 3) & (J < nj - 3)).ravel() & (cy > 1.2)
    got = (cons(q1) - cons(q0))[:, m] / dt
    return got, rhs(fields, cx[m], cy[m]), cons(q0)[:, m]


def test_residual_converges_to_axisymmetric_mhd(oracle_lib):
    errs = []
    for n in (40, 80):
        got, want, _ = residual(oracle_lib, smooth, n, 2 * n)
        scale = np.abs(want).max(axis=1)
        errs.append(np.abs(got - want).max(axis=1))
    # first order: every variable's error halves (at least by 0.6) with the grid spacing,
    # and on the finer grid it is a few per cent of that variable's right-hand side
    assert np.all(errs[1] <= 0.6 * errs[0])
    assert np.all(errs[1] <= 0.05 * scale)


@pytest.mark.parametrize("fields,var", [(rigid_rotation, 2), (pinch, 2)])
def test_equilibria_steady_to_truncation_order(oracle_lib, fields, var):
    """The radial-momentum residual of the equilibrium shrinks with the grid and is a
    small fraction of the hoop term it balances (rho w^2, or p_T - B_phi^2/4pi)."""
    errs = []
    for n in (40, 80):
        got, want, _ = residual(oracle_lib, fields, n, 2 * n)
        assert np.abs(want).max() < 1e-6           # steady: the exact right-hand side vanishes
        x = np.zeros(1)
        q = fields(x + 0.0, x + 3.0)
        hoop = abs(geometric_source(q)[2][0]) / 3.0
        errs.append(np.abs(got[var]).max())
        assert errs[-1] <= 0.02 * hoop
    assert errs[1] <= 0.6 * errs[0]


def test_swirl_and_azimuthal_field_sign_symmetry(oracle_lib):
    """(w, B_phi) -> (-w, -B_phi) in the initial state and the inflow flips them in the
    solution bit for bit and leaves every other variable unchanged (moving grid)."""
    from gelios_2d_amd.configs import mhd_heliosphere
    w = mhd_heliosphere(30, 24, 30, seed=3)
    cx, cy = w.grid.cell_centers()
    q = w.prim0.copy()
    q[4] = 0.3 * np.sin(0.2 * cx) * cy / (1 + cy)
    q[7] = 0.02 * np.cos(0.1 * cx) * cy / (1 + cy)
    outs = []
    for sgn in (1.0, -1.0):
        prm = w.params.copy()
        prm.inf_w, prm.inf_bz = sgn * 0.05, sgn * 0.01
        qq = q.copy()
        qq[4] *= sgn
        qq[7] *= sgn
        amp, period, phase = w.motion
        outs.append(oracle_lib.OracleSweep.from_polar(w.grid, prm).run_motion(qq, w.steps, w.grid.x, w.grid.y,
                                                                               amp, period, phase))
    a, b = outs
    assert np.abs(a[4]).max() > 1e-3 and np.abs(a[7]).max() > 1e-4
    assert np.array_equal(a[[0, 1, 2, 3, 5, 6, 8]], b[[0, 1, 2, 3, 5, 6, 8]])
    assert np.array_equal(a[[4, 7]], -b[[4, 7]])
