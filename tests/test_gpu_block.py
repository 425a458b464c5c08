"""The general-mesh sweep on the reference's own block grid: the grid and state read
from the reference's checkpoint (Setka::Save_Setka_ALL_ALPHA of Setka(N1..M4) filled by
Init_conditions), stepped on the device, against the reference's Go_5_komponent__MK2 on
the same file (tests/golden/sweep_block.npz)."""
import numpy as np
import pytest

from gelios_2d_amd.mesh import mesh_from_alpha, state_from_alpha
from gelios_2d_amd.params import GEL_FLUX_HLLC, reference_params
from gelios_2d_amd.state import rel_err_conserved
from test_oracle_block import GRIDS, alpha, block

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("tag", GRIDS)
def test_block_grid_sweep_matches_reference(tag):
    """The first step moves the nodes to their rails' spacing (next positions as
    Move_Setka_Calculate_2 leaves them, gel_step_moving), the others are static
    (gel_step); fitted-surface faces take the exact solver, C_1 cells the radial
    reconstruction, the C_centr wedges stay frozen."""
    from gelios_2d_amd.setka import Setka
    d = block(tag)
    a = alpha(d)
    m = mesh_from_alpha(a)
    s = Setka.from_mesh(m, reference_params())
    s.upload(state_from_alpha(a))
    s.step_moving(d["x_after"], d["y_after"])
    st = s.Go_5_komponent__MK2(int(d["steps"]) - 1)
    got = s.download()
    err = rel_err_conserved(got, d["prim"])
    x, y = s.get_nodes()
    print(f"block grid {tag} ({m.ncells} cells, {int(d['steps'])} steps): max rel err vs reference {err:.3e}, "
          f"dt {st.dt_next:.6e}, numeric {st.n_numeric}", flush=True)
    assert np.array_equal(x, d["x_after"]) and np.array_equal(y, d["y_after"])
    assert err <= TOL
    s.close()


@pytest.mark.parametrize("flux", [0, GEL_FLUX_HLLC])
def test_block_grid_matches_oracle(oracle_lib, flux):
    """HLL and HLLC (with the nul_potok faces beside the fitted surfaces) on the block
    grid against the oracle, which solves every face from both sides."""
    from gelios_2d_amd.setka import Setka
    from test_oracle_block import run_oracle
    d = block("b")
    a = alpha(d)
    m = mesh_from_alpha(a)
    prm = reference_params()
    prm.flux = flux
    steps = 25
    sw = oracle_lib.OracleSweep(m.general(), d["x"], d["y"], prm)
    sw.ptype = np.ascontiguousarray(m.ptype, np.int32)
    sw.set_nodes(d["x"], d["y"], d["x_after"], d["y_after"])
    cur, _ = sw.step(d["prim0"])
    sw.set_nodes(d["x_after"], d["y_after"], d["x_after"], d["y_after"])
    cur = sw.run(cur, steps - 1)
    s = Setka.from_mesh(m, prm)
    s.upload(d["prim0"])
    s.step_moving(d["x_after"], d["y_after"])
    s.Go_5_komponent__MK2(steps - 1)
    assert rel_err_conserved(s.download(), cur) <= TOL
    s.close()


def test_mesh_rejects_unpaired_faces():
    """An interior face whose other cell does not list its copy is an error."""
    from gelios_2d_amd._lib import GelError
    from gelios_2d_amd.setka import Setka
    d = block("a")
    m = mesh_from_alpha(alpha(d))
    k = int(np.flatnonzero(m.ftype == 0)[5])
    m.fdown[k] = (m.fdown[k] + 1) % m.ncells
    with pytest.raises(GelError):
        Setka.from_mesh(m, reference_params())


@pytest.mark.parametrize("moving", [False, True])
def test_mesh_path_equals_structured_sweep(moving):
    """The 200 x 100 heliosphere with C_1 wind rows (HLLC) given as a general mesh: the
    general-mesh kernels and the structured marching kernel agree."""
    from gelios_2d_amd.configs import heliosphere
    from gelios_2d_amd.mesh import mesh_from_polar
    from gelios_2d_amd.setka import Setka
    w = heliosphere(200, 100, 60, seed=1, flux=GEL_FLUX_HLLC)
    w.grid.cell_type[: 3 * w.grid.nj] = 0                     # three rows of C_1 wind cells
    a = Setka(w.grid, w.params)
    b = Setka.from_mesh(mesh_from_polar(w.grid), w.params)
    for s in (a, b):
        s.upload(w.prim0)
    if moving:
        rng = np.random.default_rng(3)
        for _ in range(w.steps):
            x2 = w.grid.x * (1 + 1e-4 * rng.standard_normal())
            y2 = w.grid.y * (1 + 1e-4 * rng.standard_normal())
            a.step_moving(x2, y2)
            b.step_moving(x2, y2)
    else:
        a.Go_5_komponent__MK2(w.steps)
        b.Go_5_komponent__MK2(w.steps)
    sa, sb = a.sync(), b.sync()
    err = rel_err_conserved(b.download(), a.download())
    print(f"general mesh vs structured, moving {moving}: {err:.3e}; dt {sb.dt_next:.9e} / {sa.dt_next:.9e}", flush=True)
    assert err <= TOL
    a.close()
    b.close()


def _fp32_vs_fp64(m, prm, prim, steps):
    from gelios_2d_amd.params import GEL_PREC_FP32_BULK
    from gelios_2d_amd.setka import Setka
    from gelios_2d_amd.state import l1_rel, linf_rel
    p32 = prm.copy()
    p32.precision = GEL_PREC_FP32_BULK
    a, b = Setka.from_mesh(m, prm), Setka.from_mesh(m, p32)
    for s in (a, b):
        s.upload(prim)
        s.Go_5_komponent__MK2(steps)
    sa, sb = a.sync(), b.sync()
    ra, rb = a.download(), b.download()
    a.close()
    b.close()
    assert np.all(np.isfinite(rb))
    l1 = [l1_rel(rb, ra, k) for k in (0, 1)]
    li = [linf_rel(rb, ra, k) for k in (0, 1)]
    print(f"FP32 bulk vs FP64 on the mesh, {steps} steps: L1 rho {l1[0]:.2e} p {l1[1]:.2e}; "
          f"Linf {li[0]:.2e} {li[1]:.2e}; dt {sb.dt_next:.9e} / {sa.dt_next:.9e}", flush=True)
    return l1, sa, sb


def test_mesh_fp32_bulk_heliosphere():
    """FP32 bulk on a general mesh (the 200 x 100 heliosphere with C_1 wind rows, HLLC):
    FP32 state and Riemann solver, FP64 geometry, accumulation and update, within the
    FP32-bulk tolerance of the FP64 mesh sweep (L1 <= 1e-4 per variable)."""
    from gelios_2d_amd.configs import heliosphere
    from gelios_2d_amd.mesh import mesh_from_polar
    w = heliosphere(200, 100, 200, seed=1, flux=GEL_FLUX_HLLC)
    w.grid.cell_type[: 3 * w.grid.nj] = 0
    l1, sa, sb = _fp32_vs_fp64(mesh_from_polar(w.grid), w.params, w.prim0, w.steps)
    assert max(l1) <= 1e-4
    assert abs(sb.dt_next / sa.dt_next - 1.0) < 1e-4
    assert sb.n_numeric == 0


def test_mesh_fp32_bulk_block_grid():
    """FP32 bulk on the reference's own block grid (the fitted-surface faces keep the
    FP64 exact solver, on the FP32-stored state): within L1 1e-4 of the FP64 mesh sweep."""
    d = block("b")
    a = alpha(d)
    m = mesh_from_alpha(a)
    l1, sa, sb = _fp32_vs_fp64(m, reference_params(), state_from_alpha(a), 200)
    assert max(l1) <= 1e-4
    assert abs(sb.dt_next / sa.dt_next - 1.0) < 1e-4


def test_mesh_source_table_equals_structured_sweep():
    """A per-cell source table (GEL_SRC_TABLE: mass, momentum x and y, energy, tracer, as
    the learned source leaves it) on the 200 x 100 heliosphere given as a general mesh:
    the mesh sweep equals the structured sweep given the same table."""
    from gelios_2d_amd.configs import heliosphere
    from gelios_2d_amd.mesh import mesh_from_polar
    from gelios_2d_amd.params import GEL_SRC_TABLE
    from gelios_2d_amd.setka import Setka
    w = heliosphere(200, 100, 60, seed=1, flux=GEL_FLUX_HLLC)
    w.grid.cell_type[: 3 * w.grid.nj] = 0
    prm = w.params.copy()
    prm.source_mode = GEL_SRC_TABLE
    q = 1e-3 * np.random.default_rng(4).standard_normal((5, w.grid.ncells))
    a, b = Setka(w.grid, prm), Setka.from_mesh(mesh_from_polar(w.grid), prm)
    for s in (a, b):
        s.upload(w.prim0)
        s.set_source_table(q)
        s.Go_5_komponent__MK2(w.steps)
    ra, rb = a.download(), b.download()
    np.testing.assert_array_equal(b.get_source_table(), a.get_source_table())
    a.close()
    b.close()
    err = rel_err_conserved(rb, ra)
    print(f"source table, general mesh vs structured: {err:.3e}", flush=True)
    assert np.abs(rb - w.prim0).max() > 0.0
    assert err <= TOL


def _tracked_block(tag):
    from conftest import golden
    from gelios_2d_amd.mesh import MeshRails, mesh_from_slots
    d = golden("sweep_block_tracked.npz")
    g = lambda k: d[f"{tag}_{k}"]   # noqa: E731
    m = mesh_from_slots(g("x"), g("y"), g("ptype"), g("ncorner"), g("corners"), g("cell_type"), g("nface"),
                        g("slots"))
    rails = MeshRails(g("rail_kind"), g("rail_m"), g("rail_s"), g("rail_offset"), g("rail_all"), g("rail_key"),
                      (g("line_contact"), g("line_inner"), g("line_outer")))
    return d, g, m, rails


@pytest.mark.parametrize("tag", GRIDS)
def test_block_grid_tracked_matches_reference(tag):
    """The reference's block grid with its fitted surfaces moving: every step runs
    Move_surface over the three lines and Move_Setka_Calculate_2 over the A, B, C and D
    rails on the device, then the sweep from the nodes before to those after, against
    the reference's Go_5_komponent__MK2 for 1 step and for 20 (state 1e-12 relative per
    conserved variable; nodes 1e-12 of the largest coordinate: the device's pow, cos
    and atan can differ from glibc's in the last bit)."""
    from gelios_2d_amd.params import N_OUTER_SHOCK, reference_tracking
    from gelios_2d_amd.setka import Setka
    d, g, m, rails = _tracked_block(tag)
    N = g("N")
    t = reference_tracking(outer_faces=N_OUTER_SHOCK + int(N[0]) + int(N[1]) + 1)
    for steps, sfx in ((1, "1"), (int(g("steps")), "n")):
        s = Setka.from_mesh(m, reference_params())
        s.upload(g("prim0"))
        s.set_mesh_tracking(t, rails)
        st = s.Go_5_komponent__MK2(steps)
        got = s.download()
        x, y = s.get_nodes()
        xr, yr = g("x_" + sfx), g("y_" + sfx)
        nerr = max(np.abs(x - xr).max(), np.abs(y - yr).max()) / max(np.abs(xr).max(), np.abs(yr).max())
        err = rel_err_conserved(got, g("prim_" + sfx))
        print(f"tracked block grid {tag} ({m.ncells} cells), {steps} steps: state {err:.3e}, nodes {nerr:.3e}, "
              f"numeric {st.n_numeric}", flush=True)
        assert nerr <= TOL
        assert err <= TOL
        s.close()


def test_mesh_tracking_rejects_bad_tables():
    """Lines whose faces are not faces of the mesh, and shock lines longer than the
    contact line (the reference's averaging loops index the contact line by the
    shocks' face counts), are refused."""
    from gelios_2d_amd._lib import GelError
    from gelios_2d_amd.mesh import MeshRails
    from gelios_2d_amd.params import reference_tracking
    from gelios_2d_amd.setka import Setka
    d, g, m, rails = _tracked_block("b")
    s = Setka.from_mesh(m, reference_params())
    s.upload(g("prim0"))
    bad = g("line_contact").copy()
    bad[0, 2], bad[0, 3] = bad[0, 3] + 1, bad[0, 2] + 1
    r2 = MeshRails(rails.kind, rails.m, rails.s, rails.all_offset, rails.all, rails.key,
                   (bad, rails.lines[1], rails.lines[2]))
    with pytest.raises(GelError):
        s.set_mesh_tracking(reference_tracking(), r2)
    r3 = MeshRails(rails.kind, rails.m, rails.s, rails.all_offset, rails.all, rails.key,
                   (rails.lines[1], rails.lines[0], rails.lines[2]))
    with pytest.raises(GelError):
        s.set_mesh_tracking(reference_tracking(), r3)
    s.close()
