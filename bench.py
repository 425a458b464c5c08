"""Benchmark of the Godunov flux-and-update sweep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one pass of the sweep over the whole grid (one launch of the sweep
kernel); a cell-update is one cell advanced by one step.  Inputs are resident in
HBM when the timed region starts.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s on 1 MI355X: FP32+surrogate vs FP64 plain at L1 err ≤1e-4"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from gelios_2d_amd import configs
    if name == "cfg1":
        return configs.config1()
    if name == "gas800x400":
        return configs.heliosphere(800, 400, 1000, seed=2, name="hllc-gas-fp64-800x400")
    if name.startswith("large"):
        return configs.large()
    raise SystemExit(f"unknown workload {name}")


def bytes_per_cell_update(w):
    # state read once and written once (5 FP64 primitives), node x,y read (16 B)
    return 2 * 5 * 8 + 16


def cpu_baseline(w, budget_s=12.0):
    """The CPU restatement of the reference sweep (per-cell gather over faces, every
    interior face solved from both sides, FP64, OpenMP over cells, as
    Setka.cpp:10148) on the host cores, over a bounded number of steps."""
    from oracle import oracle as ora
    threads = min(16, os.cpu_count() or 1)
    sw = ora.OracleSweep.from_polar(w.grid, w.params)
    cur = w.prim0.copy()
    cur, _ = sw.step(cur, nthreads=threads)            # warm-up
    t0 = time.perf_counter()
    cur, _ = sw.step(cur, nthreads=threads)
    one = time.perf_counter() - t0
    steps = max(1, min(w.steps, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(steps):
        cur, _ = sw.step(cur, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": w.grid.ncells * steps / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of {w.name} ({w.grid.ncells} cells), OpenMP {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="gas800x400")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    from gelios_2d_amd.setka import Setka
    w = workload(args.workload)
    K = args.steps if args.steps is not None else w.steps
    W = args.warmup
    nc = w.grid.ncells
    log(f"bench: workload {w.name}, {nc} cells, warmup {W}, steps {K}")

    s = Setka(w.grid, w.params)
    s.upload(w.prim0)
    s.Go_5_komponent__MK2(W)
    s.sync()
    t0 = time.perf_counter()
    s.Go_5_komponent__MK2(K)
    st = s.sync()
    wall = time.perf_counter() - t0
    value = nc * K / wall
    log(f"bench: {K} steps in {wall:.4f} s -> {value:.4e} cell-updates/s, dt_next {st.dt_next:.3e}, "
        f"clamps rho {st.n_rho_clamp} p {st.n_p_clamp}")

    # per-launch device time of the sweep kernel (HIP events on the context's stream)
    s.enable_timing(True)
    s.Go_5_komponent__MK2(K)
    avg_ms, n = s.kernel_timing()
    s.enable_timing(False)
    B = bytes_per_cell_update(w)
    achieved = B * nc / (avg_ms * 1e-3) / 1e9
    log(f"bench: sweep kernel avg {avg_ms * 1e3:.2f} us over {n} launches, {achieved:.1f} GB/s algorithmic")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w.name, "cells": nc, "grid": [w.grid.ni, w.grid.nj], "flux": "HLLC",
                   "precision": "fp64-plain", "cfl": w.params.cfl},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_cell_update": B, "kernel_avg_us": avg_ms * 1e3},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w)
        log(f"bench: cpu baseline {line['cpu_baseline']['value']:.3e} cell-updates/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
