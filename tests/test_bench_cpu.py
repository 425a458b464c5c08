"""bench.py's contract pieces that need no GPU: the algorithmic bytes per cell-update
behind roofline.achieved, and the one-GPU limit."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_algorithmic_bytes_per_cell_update():
    """2 x variables x bytes per value (state read and written) + node positions (16 B
    static, 32 B moving) + per-cell source inputs in the state's precision."""
    import bench
    from gelios_2d_amd import configs
    from gelios_2d_amd.params import GEL_PREC_FP32_BULK, GEL_SRC_HYBRID
    assert bench.bytes_per_cell_update(configs.hllc_gas_200x100()) == 2 * 5 * 8 + 16
    assert bench.bytes_per_cell_update(configs.hlld_mhd_800x400()) == 2 * 9 * 8 + 32
    assert bench.bytes_per_cell_update(configs.hlld_mhd_800x400(GEL_PREC_FP32_BULK)) == 2 * 9 * 4 + 32
    w = configs.mhd_heliosphere(20, 10, 1, seed=1)
    w.params.source_mode = 1                         # GEL_SRC_MK1: 19 hydrogen fields
    assert bench.bytes_per_cell_update(w) == 2 * 9 * 8 + 32 + 19 * 8
    w = configs.mhd_heliosphere(20, 10, 1, seed=1, moving=False, precision=GEL_PREC_FP32_BULK)
    w.params.source_mode = GEL_SRC_HYBRID            # the ensemble: flag + 24 inputs in FP32
    assert bench.bytes_per_cell_update(w) == 2 * 9 * 4 + 16 + 4 + 24 * 4


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
