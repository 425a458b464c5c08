"""The reference's own block grid (Setka(N1..M4), Setka.cpp:19-1076: rails of four kinds,
fitted-surface rows, triangular C_centr wedges): its checkpoint read into face tables,
and the oracle's sweep on it, against the compiled reference (tests/golden/sweep_block.npz,
written by tests/golden/make_golden.py through oracle/ref_harness.cpp, mode_block)."""
import os
import tempfile

import numpy as np
import pytest

from conftest import golden
from gelios_2d_amd.checkpoint import read_alpha
from gelios_2d_amd.grid import MAXF
from gelios_2d_amd.mesh import mesh_from_alpha, state_from_alpha
from gelios_2d_amd.params import reference_params

GRIDS = ("a", "b")


def block(tag):
    d = golden("sweep_block.npz")
    return {k[2:]: d[k] for k in d.files if k.startswith(tag + "_")}


def alpha(d):
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "ck.txt")
        with open(p, "wb") as f:
            f.write(d["ref_text"].tobytes())
        return read_alpha(p)


@pytest.mark.parametrize("tag", GRIDS)
def test_checkpoint_gives_the_reference_face_tables(tag):
    """Download_Setka_ALL_ALPHA_2_0 + TVD_prepare restated (mesh_from_alpha) give the
    node positions, point types, contours, cell types and every face slot (A, B, type,
    Sosed, Sosed_down, Sosed_up) the reference holds after the same two calls."""
    d = block(tag)
    a = alpha(d)
    m = mesh_from_alpha(a)
    assert np.array_equal(m.x, d["x"]) and np.array_equal(m.y, d["y"])
    assert np.array_equal(m.ptype, d["ptype"])
    assert np.array_equal(np.diff(m.coff), d["ncorner"]) and np.array_equal(m.corners, d["corners"])
    assert np.array_equal(m.cell_type, d["cell_type"]) and np.array_equal(m.nface, d["nface"])
    ref = d["slots"]
    for k, name in enumerate(("fa", "fb", "ftype", "fnbr", "fdown", "fup")):
        assert np.array_equal(getattr(m, name).reshape(-1, MAXF), ref[:, :, k]), name
    assert np.array_equal(state_from_alpha(a), d["prim0"])
    assert (m.ptype > 0).sum() > 30 and (np.diff(m.coff) == 3).sum() > 0


def run_oracle(oracle_lib, d, m, steps):
    """The reference's step loop on this grid: the first step moves the nodes to their
    rails' spacing (x -> x_after), the later ones leave them there."""
    sw = oracle_lib.OracleSweep(m.general(), d["x"], d["y"], reference_params())
    sw.ptype = np.ascontiguousarray(m.ptype, np.int32)
    sw.set_nodes(d["x"], d["y"], d["x_after"], d["y_after"])
    cur, st = sw.step(d["prim0"], nthreads=1)
    sw.set_nodes(d["x_after"], d["y_after"], d["x_after"], d["y_after"])
    for _ in range(steps - 1):
        cur, st = sw.step(cur, nthreads=1)
    return cur


@pytest.mark.parametrize("tag", GRIDS)
def test_oracle_matches_reference_on_block_grid(oracle_lib, tag):
    """The oracle's cell loop (solving every face from both sides, as the reference) is
    bit for bit the reference's Go_5_komponent__MK2 on its own grid."""
    d = block(tag)
    m = mesh_from_alpha(alpha(d))
    got = run_oracle(oracle_lib, d, m, int(d["steps"]))
    assert np.array_equal(got, d["prim"])


def tracked_block(tag):
    """The tracked block-grid fixture (tests/golden/sweep_block_tracked.npz, harness mode
    blocktrack) as (getter, mesh, rails, tracking constants)."""
    from gelios_2d_amd.mesh import MeshRails, mesh_from_slots
    from gelios_2d_amd.params import N_OUTER_SHOCK, reference_tracking
    d = golden("sweep_block_tracked.npz")
    g = lambda k: d[f"{tag}_{k}"]   # noqa: E731
    m = mesh_from_slots(g("x"), g("y"), g("ptype"), g("ncorner"), g("corners"), g("cell_type"), g("nface"),
                        g("slots"))
    rails = MeshRails(g("rail_kind"), g("rail_m"), g("rail_s"), g("rail_offset"), g("rail_all"), g("rail_key"),
                      (g("line_contact"), g("line_inner"), g("line_outer")))
    N = g("N")
    t = reference_tracking(outer_faces=N_OUTER_SHOCK + int(N[0]) + int(N[1]) + 1)
    return g, m, rails, t


@pytest.mark.parametrize("tag", GRIDS)
def test_tracked_block_node_motion_matches_reference(tag):
    """Move_surface over the three lines and Move_Setka_Calculate_2 over the A, B, C, D
    rails (oracle/mesh_tracking.py) give the reference's next node positions bit for
    bit from the fixture's first step (dt = 1e-8, the state before it)."""
    from oracle import mesh_tracking as mt
    g, m, rails, t = tracked_block(tag)
    x2, y2 = mt.move_nodes(m, rails, t, g("x"), g("y"), g("prim0"), 1e-8, 5.0 / 3.0)
    assert np.array_equal(x2, g("x_1")) and np.array_equal(y2, g("y_1"))


@pytest.mark.parametrize("tag", GRIDS)
def test_tracked_block_oracle_matches_reference(oracle_lib, tag):
    """The oracle's tracked steps (node motion, cell loop, node commit) against the
    reference's Go_5_komponent__MK2 over 1 and 20 steps: nodes bit for bit, state
    within 1e-13 relative (one ulp of drift in one cell over 20 steps)."""
    from gelios_2d_amd.state import rel_err_conserved
    from oracle import mesh_tracking as mt
    g, m, rails, t = tracked_block(tag)
    for steps, sfx in ((1, "1"), (int(g("steps")), "n")):
        o = mt.MeshTrackedOracle(m, rails, t, reference_params())
        p = o.run(g("prim0"), steps)
        assert np.array_equal(o.x, g("x_" + sfx)) and np.array_equal(o.y, g("y_" + sfx))
        assert rel_err_conserved(p, g("prim_" + sfx)) <= 1e-13
