"""Fairness check of the CPU baseline: the compiled reference sweep against the oracle.

Test infrastructure, run only in the build container, where /root/reference exists
(oracle/Makefile builds oracle/_ref/ref_harness from the reference sources).  Both
sides advance the same polar grid (C_1 wind rows at the inner boundary, the
reference's constants: HLL, CFL 0.1) from the same state for the same number of steps
with the same OpenMP thread count; only the step loop is timed (the harness times its
call of Setka::Go_5_komponent__MK2 with std::chrono::steady_clock).

    python oracle/time_reference.py [--ni 200] [--nj 100] [--steps 20] [--threads 8] [--write]

Prints the cell-updates/s of both and their ratio; --write adds the measurement to
oracle/fairness.json (one entry per thread count), which the CPU-baseline leg of the
benchmark reads on hosts where the reference is absent.  With more than one thread the
reference's shared time step races (Setka.cpp:10141), so only one thread is compared
bit for bit.
"""
from __future__ import annotations

import argparse
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

from gelios_2d_amd.params import reference_params  # noqa: E402
from oracle import oracle as ora  # noqa: E402


REF_SOURCE = "/root/reference/Setka.cpp"


def available() -> bool:
    """The reference sources are present (the build container) and the harness is
    built from them.  Elsewhere (a GPU host) the reference is never run, even where a
    stale harness binary exists."""
    return os.path.exists(REF_SOURCE) and os.path.exists(ora.REF_HARNESS)


def measure(ni: int = 200, nj: int = 100, steps: int = 20, threads: int = 1) -> dict:
    """Time both sweeps on the same grid, state, steps and threads; cell-updates/s."""
    import make_golden as mg
    if not available():
        raise RuntimeError("the reference harness is not built (needs the reference sources)")
    prm = reference_params()
    g = mg.harness_grid(ni, nj, n_c1=max(ni // 5, 1))
    prim0 = mg.initial_state(g, np.random.default_rng(1), prm)
    nc = g.ncells

    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "sweep.in"), os.path.join(tmp, "sweep.out")
        with open(fin, "wb") as f:
            f.write(mg.sweep_blob(g, steps, prim0))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([ora.REF_HARNESS, "sweep", fin, fout], check=True, cwd=tmp, env=env,
                           stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
        t_ref = float(re.search(r"ref_sweep_seconds ([0-9.eE+-]+)", r.stderr).group(1))
        with open(fout, "rb") as f:
            ref_prim = mg.parse_sweep(f.read(), g)[0]

    sw = ora.OracleSweep.from_polar(g, prm)
    sw.step(prim0, nthreads=threads)               # warm-up (threads, pages)
    sw = ora.OracleSweep.from_polar(g, prm)
    t0 = time.perf_counter()
    out = sw.run(prim0, steps, nthreads=threads)
    t_ora = time.perf_counter() - t0
    diff = float(np.max(np.abs(out - ref_prim) / np.maximum(np.abs(ref_prim), 1e-300)))
    return {"grid": [ni, nj], "steps": steps, "threads": threads, "reference": nc * steps / t_ref,
            "port": nc * steps / t_ora, "max_rel_diff": diff}


FAIRNESS_JSON = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fairness.json")


def write_fairness(m: dict) -> None:
    import json
    import platform
    d = {"entries": []}
    if os.path.exists(FAIRNESS_JSON):
        with open(FAIRNESS_JSON) as f:
            d = json.load(f)
    cpu = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        cpu = next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), cpu)
    except OSError:
        pass
    e = dict(m, port_over_reference=m["port"] / m["reference"], cpu=cpu,
             workload="the reference's gas HLL sweep (Go_5_komponent__MK2), polar grid with C_1 wind rows")
    d["entries"] = [x for x in d["entries"] if x["threads"] != m["threads"]] + [e]
    d["entries"].sort(key=lambda x: x["threads"])
    with open(FAIRNESS_JSON, "w") as f:
        json.dump(d, f, indent=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ni", type=int, default=200)
    ap.add_argument("--nj", type=int, default=100)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--write", action="store_true")
    args = ap.parse_args()
    ora.build()
    m = measure(args.ni, args.nj, args.steps, args.threads)
    print(f"grid {args.ni}x{args.nj}, {args.steps} steps, {args.threads} threads")
    print(f"reference  {m['reference']:.4e} cell-updates/s")
    print(f"oracle     {m['port']:.4e} cell-updates/s")
    print(f"oracle / reference speed: {m['port'] / m['reference']:.2f}; "
          f"max relative difference of the states {m['max_rel_diff']:.2e}")
    if args.write:
        write_fairness(m)


if __name__ == "__main__":
    main()
