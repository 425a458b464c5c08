"""Benchmark of the Godunov flux-and-update sweep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one pass of the sweep over the whole grid (one launch of the sweep
kernel); a cell-update is one cell advanced by one step.  Inputs are resident in
HBM when the timed region starts.  Prints one JSON line.

Workloads
  cfg2 (default)   HLLD MHD, 9 variables, 800 x 400 cells, moving grid, 1000 steps:
                   the FP32-bulk sweep (value) beside the FP64 plain sweep, with the
                   density / pressure L1 error of the former against the latter
  cfg1             single-fluid HLLC, 200 x 100 cells, 100 steps, FP64
  large-mhd        cfg2 physics on 4096 x 2048 cells (state far beyond the 256 MB
                   Infinity Cache): the HBM roofline of the FP64 sweep
  large-gas        cfg1 physics on 4096 x 2048 cells, FP64
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


def bytes_per_cell_update(w):
    """Algorithmic HBM bytes per cell-update: state read once and written once, node
    coordinates read (16 B), plus 16 B of written node coordinates on a moving grid
    (the per-unit figures of the survey's measurement table)."""
    from gelios_2d_amd.params import GEL_PREC_FP32_BULK
    nv = 9 if w.params.physics == 1 else 5
    s = 4 if w.params.precision == GEL_PREC_FP32_BULK else 8
    geom = 32 if w.motion else 16
    return 2 * nv * s + geom


def run_gpu(w, K, W, label):
    from gelios_2d_amd.setka import Setka
    nc = w.grid.ncells
    s = Setka(w.grid, w.params)
    if w.motion:
        s.set_motion(*w.motion)
    s.upload(w.prim0)
    s.Go_5_komponent__MK2(W)
    s.sync()
    t0 = time.perf_counter()
    s.Go_5_komponent__MK2(K)
    st = s.sync()
    wall = time.perf_counter() - t0
    value = nc * K / wall
    log(f"bench[{label}]: {K} steps in {wall:.4f} s -> {value:.4e} cell-updates/s, dt_next {st.dt_next:.3e}, "
        f"clamps rho {st.n_rho_clamp} p {st.n_p_clamp}")
    final = s.download()
    # per-launch device time of the sweep kernel (HIP events on the context's stream)
    s.enable_timing(True)
    s.Go_5_komponent__MK2(min(K, 200))
    avg_ms, n = s.kernel_timing()
    s.enable_timing(False)
    B = bytes_per_cell_update(w)
    achieved = B * nc / (avg_ms * 1e-3) / 1e9
    log(f"bench[{label}]: sweep kernel avg {avg_ms * 1e3:.2f} us over {n} launches, "
        f"{achieved:.1f} GB/s algorithmic ({B} B/cell-update)")
    s.close()
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_cell_update": B,
            "kernel_avg_us": avg_ms * 1e3}
    return {"value": value, "wall_s": wall, "final": final, "roofline": roof, "steps_done": K + W,
            "dt_next": st.dt_next}


def cpu_baseline(w, budget_s=12.0):
    """The CPU restatement of the reference sweep (per-cell gather over faces, every
    interior face solved from both sides, FP64, OpenMP over cells, as
    Setka.cpp:10148) on the host cores, over a bounded number of steps."""
    from oracle import oracle as ora
    threads = min(16, os.cpu_count() or 1)
    sw = ora.OracleSweep.from_polar(w.grid, w.params)
    x0, y0 = w.grid.x, w.grid.y
    amp, period, phase = w.motion if w.motion else (0.0, 1, 0.0)

    def one(cur, n):
        if w.motion:
            from gelios_2d_amd.configs import motion_scales
            sc = motion_scales(amp, period, phase)
            so, sn = sc[n % period], sc[(n + 1) % period]
            sw.set_nodes(x0 * so, y0 * so, x0 * sn, y0 * sn)
        nxt, _ = sw.step(cur, nthreads=threads)
        return nxt

    cur = one(w.prim0.copy(), 0)            # warm-up
    t0 = time.perf_counter()
    cur = one(cur, 1)
    t1 = time.perf_counter() - t0
    steps = max(1, min(w.steps, int(budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    for n in range(steps):
        cur = one(cur, 2 + n)
    dt = time.perf_counter() - t0
    return {"value": w.grid.ncells * steps / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of {w.name} ({w.grid.ncells} cells), FP64, OpenMP {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    from gelios_2d_amd import configs
    from gelios_2d_amd.params import GEL_PREC_FP32_BULK, GEL_PREC_FP64
    from gelios_2d_amd.state import l1_rel

    wl = args.workload
    W = args.warmup
    extra = {}
    if wl == "cfg2":
        w64, w32 = configs.config2(GEL_PREC_FP64), configs.config2(GEL_PREC_FP32_BULK)
        K = args.steps if args.steps is not None else w64.steps
        log(f"bench: {w64.name} / {w32.name}, {w64.grid.ncells} cells, warmup {W}, steps {K}")
        r64 = run_gpu(w64, K, W, "fp64-plain")
        r32 = run_gpu(w32, K, W, "fp32-bulk")
        main_w, main_r = w32, r32
        extra = {
            "fp64_plain": {"value": r64["value"], "ms_per_step": r64["wall_s"] * 1e3 / K,
                           "roofline": r64["roofline"]},
            "speedup_fp32_vs_fp64": r32["value"] / r64["value"],
            "l1_vs_fp64": {"density": l1_rel(r32["final"], r64["final"], 0),
                           "pressure": l1_rel(r32["final"], r64["final"], 1),
                           "after_steps": W + K},
            "surrogate": "not in this line: cfg2 carries no charge-exchange sources",
        }
        precision, dtype = "fp32-bulk (FP32 state and face arithmetic; FP64 geometry, accumulation, update)", "f32"
    else:
        if wl == "cfg1":
            w = configs.config1()
        elif wl == "large-mhd":
            w = configs.mhd_heliosphere(4096, 2048, 20, seed=2, name="hlld-fp64-4096x2048-moving")
        elif wl == "large-gas":
            w = configs.large()
        else:
            raise SystemExit(f"unknown workload {wl}")
        K = args.steps if args.steps is not None else w.steps
        log(f"bench: {w.name}, {w.grid.ncells} cells, warmup {W}, steps {K}")
        main_w, main_r = w, run_gpu(w, K, W, "fp64-plain")
        precision, dtype = "fp64-plain", "f64"

    line = {
        "metric": METRIC,
        "value": main_r["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": main_r["wall_s"] * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": main_w.name, "cells": main_w.grid.ncells, "grid": [main_w.grid.ni, main_w.grid.nj],
                   "physics": "mhd-hlld" if main_w.params.physics == 1 else
                   ("gas-hllc" if main_w.params.flux == 1 else "gas-hll"),
                   "moving_grid": bool(main_w.motion), "precision": precision, "cfl": main_w.params.cfl},
        "roofline": main_r["roofline"],
    }
    line.update(extra)
    if not args.no_cpu:
        cw = main_w if wl != "cfg2" else w64
        line["cpu_baseline"] = cpu_baseline(cw)
        log(f"bench: cpu baseline {line['cpu_baseline']['value']:.3e} cell-updates/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
