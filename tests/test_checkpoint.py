"""The reference's checkpoint format (Setka::Save_Setka_ALL_ALPHA, Setka.cpp:2355-2574):
the writer reproduces, byte for byte, the file the compiled reference writes for the
same polar grid and state (tests/golden/checkpoint_polar.npz), the reader recovers grid
and state to the six digits the format keeps, and a read file writes back unchanged."""
import numpy as np

from conftest import golden


def _case():
    from gelios_2d_amd.grid import PolarGrid
    d = golden("checkpoint_polar.npz")
    g = PolarGrid(int(d["ni"]), int(d["nj"]), d["x"], d["y"], d["cell_type"], d["outer_type"])
    return d, g


def test_writer_matches_reference_bytes(tmp_path):
    from gelios_2d_amd import checkpoint as ck
    d, g = _case()
    p = tmp_path / "ours.txt"
    ck.write_alpha(p, ck.alpha_from_polar(g, d["prim"], d["src"]))
    assert p.read_bytes() == d["ref_text"].tobytes()


def test_reader_recovers_grid_and_state(tmp_path):
    from gelios_2d_amd import checkpoint as ck
    d, g = _case()
    p = tmp_path / "ref.txt"
    p.write_bytes(d["ref_text"].tobytes())
    m = ck.read_alpha(p)
    g2, prim, src = ck.polar_from_alpha(m)
    assert (g2.ni, g2.nj) == (g.ni, g.nj)
    assert np.array_equal(g2.cell_type, g.cell_type) and np.array_equal(g2.outer_type, g.outer_type)
    for got, want in ((g2.x, g.x), (g2.y, g.y), (prim, d["prim"]), (src, d["src"])):
        assert np.all(np.abs(got - want) <= 5e-6 * np.abs(want) + 1e-300)
    q = tmp_path / "again.txt"
    ck.write_alpha(q, m)
    assert q.read_bytes() == p.read_bytes()


def test_non_polar_mesh_is_refused(tmp_path):
    import pytest
    from gelios_2d_amd import checkpoint as ck
    d, g = _case()
    m = ck.alpha_from_polar(g, d["prim"], d["src"])
    m.contour[3] = list(reversed(m.contour[3]))
    with pytest.raises(ValueError):
        ck.polar_from_alpha(m)
