"""Build libgelios.so in-tree with hipcc for gfx950 (no JIT cache, no pip install)."""
from __future__ import annotations

import os
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgelios.so")
SOURCES = ["gelios_sweep.hip", "sweep_fp32.hip", "surrogate.hip", "tracking.hip", "multifluid.hip", "inner.hip",
           "mesh.hip", "mc.hip"]
HIPCC = "/opt/rocm/bin/hipcc"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # parity: no FMA contraction, so device arithmetic rounds like the reference's
    # (sweep_fp32.hip turns contraction on for its own FP32-bulk kernels)
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    """Compile every source to an object (in parallel), link the shared library."""
    with tempfile.TemporaryDirectory() as tmp:
        def compile_one(src):
            obj = os.path.join(tmp, src + ".o")
            r = subprocess.run([HIPCC, *HIPCC_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)],
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"hipcc failed on {src}:\n" + r.stderr[-4000:])
            if verbose and r.stderr:
                print(r.stderr)
            return obj
        with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
            objs = list(ex.map(compile_one, SOURCES))
        tmp_lib = LIB + ".tmp"
        r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_lib, *objs],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr[-4000:])
        os.replace(tmp_lib, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=False))
