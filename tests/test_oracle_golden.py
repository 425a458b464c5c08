"""The CPU oracle against vectors written by the compiled reference (tests/golden/).

Flux, limiter helpers, cell geometry and source are compared point-wise; the whole
sweep (Setka::Go_5_komponent__MK2, Setka.cpp:10121-10482) is compared after N steps
on a polar grid built from the reference's own Point/Cell/Gran objects.  The bar is
bit-exact: the restatement keeps the reference's floating-point operation order.
"""
import numpy as np
import pytest

from conftest import golden
from gelios_2d_amd.grid import PolarGrid, pack_general
from gelios_2d_amd.params import GEL_SRC_MK1, reference_params


def test_hllc_vectors_bitexact(oracle_lib):
    d = golden("hllc_vectors.npz")
    inp, out = d["inp"], d["out"]
    for k in range(len(inp)):
        r = inp[k]
        P, PQ, Vl, Vc, Vp, t = oracle_lib.hllc2d(r[0:5], r[5:10], r[10], r[11], r[12], r[13], 0.1, 5.0 / 3.0,
                                                 int(r[14]), int(r[15]))
        got = np.array(list(P) + [PQ, Vl, Vc, Vp, t])
        assert np.array_equal(got, out[k]), (k, got, out[k])


def test_helpers_vectors(oracle_lib):
    L = oracle_lib.lib()
    d = golden("helpers_vectors.npz")
    import ctypes
    for r, o in zip(d["inp"], d["out"]):
        assert L.ora_minmod(r[0], r[1]) == o[0]
        assert L.ora_linear3(*r) == o[1]
        assert L.ora_linear2(r[0], r[1], r[2], r[3], r[6]) == o[2]
        assert L.ora_polar_angle(r[0], r[1]) == o[3]
        u, v = ctypes.c_double(r[4]), ctypes.c_double(r[5])
        L.ora_polar_perenos(r[0], r[1], r[2], r[3], ctypes.byref(u), ctypes.byref(v))
        assert (u.value, v.value) == (o[4], o[5])


def test_cell_geometry_and_source(oracle_lib):
    import ctypes
    L = oracle_lib.lib()
    prm = reference_params()
    d = golden("cell_vectors.npz")
    q = np.arange(4, dtype=np.int32)
    for r, o in zip(d["inp"], d["out"]):
        px = np.ascontiguousarray(r[0:8:2])
        py = np.ascontiguousarray(r[1:8:2])
        cx, cy = ctypes.c_double(), ctypes.c_double()
        L.ora_cell_center(oracle_lib._d(px), oracle_lib._d(py), oracle_lib._i(q), ctypes.byref(cx), ctypes.byref(cy))
        assert (cx.value, cy.value) == (o[0], o[1])
        assert L.ora_cell_area(oracle_lib._d(px), oracle_lib._d(py), oracle_lib._i(q)) == o[2]
        assert L.ora_cell_volume_rotate(oracle_lib._d(px), oracle_lib._d(py), oracle_lib._i(q), 360.0) == o[3]
        f = np.ascontiguousarray(r[12:31])
        q1, q2, q3 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        L.ora_source_mk1(ctypes.byref(prm), oracle_lib._d(f), r[8], r[9], r[10], r[11],
                         ctypes.byref(q1), ctypes.byref(q2), ctypes.byref(q3))
        assert (q1.value, q2.value, q3.value) == (o[4], o[5], o[6])


def _grid(d):
    return PolarGrid(int(d["ni"]), int(d["nj"]), d["x"], d["y"], d["cell_type"], d["outer_type"])


@pytest.mark.parametrize("name", ["sweep_static.npz", "sweep_moving.npz", "sweep_sources.npz"])
def test_stencil_tables_match_tvd_prepare(name):
    """The packer's per-cell face slots equal what Setka::TVD_prepare built."""
    d = golden(name)
    g = _grid(d)
    gm = pack_general(g)
    nc = g.ncells
    mine = np.stack([gm.fa, gm.fb, gm.ftype, gm.fnbr, gm.fdown, gm.fup], 1).reshape(nc, 8, 6)
    assert np.array_equal(mine[:, :4], d["slots"][:, :4])
    assert np.array_equal(gm.contour, d["contour"])
    assert np.all(d["nface"] == 4)


@pytest.mark.parametrize("name", ["sweep_static.npz", "sweep_moving.npz", "sweep_sources.npz"])
def test_whole_sweep_bitexact(oracle_lib, name):
    d = golden(name)
    g = _grid(d)
    prm = reference_params()
    if "src" in d.files:
        prm.source_mode = GEL_SRC_MK1
    sw = oracle_lib.OracleSweep(pack_general(g), g.x, g.y, prm)
    if "src" in d.files:
        sw.src = np.ascontiguousarray(d["src"])
    steps = int(d["steps"])
    if "traj" in d.files:
        cur = d["prim0"].copy()
        x, y = g.x.copy(), g.y.copy()
        for s in range(steps):
            x2, y2 = d["traj"][s]
            sw.set_nodes(x, y, x2, y2)
            cur, st = sw.step(cur, dt_prev=prm.dt0)
            assert st == 0
            x, y = x2, y2
        res = cur
    else:
        res = sw.run(d["prim0"], steps)
    assert np.array_equal(res, d["prim"])


def test_port_matches_compiled_reference_sweep():
    """Where the reference harness is built (the build container): on a 40 x 20 polar
    grid with C_1 wind rows the oracle's sweep reproduces the reference's own
    Go_5_komponent__MK2 bit for bit over 5 steps (one thread: the reference's shared
    time step races with more)."""
    from oracle import time_reference as tr
    if not tr.available():
        pytest.skip("reference harness not built (the reference sources are absent)")
    m = tr.measure(40, 20, 5, threads=1)
    assert m["max_rel_diff"] == 0.0
    assert m["reference"] > 0 and m["port"] > 0
