"""Benchmark of the Godunov flux-and-update sweep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one pass of the sweep over the whole grid (one launch of the sweep
kernel); a cell-update is one cell advanced by one step.  Inputs are resident in
HBM when the timed region starts.  Prints one JSON line.

Workloads
  ensemble (default)  the metric's comparison: 64 heliosphere cases (100x50 .. 800x400
                   cells, seeded inflow / field / hydrogen), FP32-bulk HLLD with the
                   learned charge-exchange source, one persistent launch per step,
                   HIP-graph replay, launch shape from the autotuner, 500 steps (value);
                   beside it the FP64 plain sweep of hlld-800x400 and the density / pressure L1
                   error of the batch against an FP64 run of the same cases with the
                   analytic sources
  hlld-800x400     HLLD MHD, 9 variables, 800 x 400 cells, moving grid, 1000 steps:
                   the FP32-bulk sweep (value) beside the FP64 plain sweep
  hllc-200x100     single-fluid HLLC, 200 x 100 cells, 100 steps, FP64
  large-mhd        hlld-800x400 physics on 4096 x 2048 cells (state far beyond the 256 MB
                   Infinity Cache): the HBM roofline of the FP64 sweep
  large-gas        hllc-200x100 physics on 4096 x 2048 cells, FP64
  tracked-800x400  a heliosphere with the three fitted surfaces tracked every step
                   (exact-solver surface speeds, A-rail redistribution, sweep with
                   exact-solver surface faces), 799 x 400 cells, FP64 HLL, the
                   reference's constants (CFL 0.1), 200 steps
  hydrogen-1600x800  HLLD MHD, 1600 x 800 cells, moving grid, charge exchange with four
                   frozen hydrogen populations, FP64 plain beside FP32 bulk, 2000 steps
  plasma-4h-800x400  the plasma and four hydrogen fluids (21 fields, HLLC, charge
                   exchange), 800 x 400 cells, static, FP64, 200 steps
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
# HBM bytes per launch of the batched ensemble's sweep kernel (hybrid source), measured
# with rocprofv3 --pmc: FETCH_SIZE (scaled x2: a 1 GiB device copy reads as 524 302 KiB)
# plus WRITE_SIZE, in KiB per dispatch (profiles/sweep_batch_ensemble_64_cases_fp32_hybrid.md)
ENSEMBLE_TRAFFIC_BYTES = (2 * 746319 + 196987) * 1024
METRIC = "cell-updates/s on 1 MI355X: FP32+surrogate vs FP64 plain at L1 err ≤1e-4"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_cell_update(w, fallback_fraction=0.0):
    """Algorithmic HBM bytes per cell-update: the state read once and written once
    (2 x variables x bytes per value), one node's coordinates (16 B; 32 B on a moving
    grid, positions before and after the step), and the per-cell source inputs, held in
    the state's precision: the table (5 values), the hydrogen moments (19 values), or
    in the hybrid mode a 4-byte flag and all of them (the static FP32-bulk sweep copies
    every input of a row into LDS), the FP64 sweep the moments for the fallback
    fraction of cells and the table for the rest."""
    from gelios_2d_amd.params import GEL_PREC_FP32_BULK, GEL_SRC_HYBRID, GEL_SRC_MK1, GEL_SRC_TABLE
    nv = {1: 9, 2: 21}.get(w.params.physics, 5)
    s = 4 if w.params.precision == GEL_PREC_FP32_BULK else 8
    moving = bool(getattr(w, "motion", None) or getattr(w, "tracking", None))
    geom = 32 if moving else 16
    src = {GEL_SRC_TABLE: 5 * s, GEL_SRC_MK1: 19 * s}.get(w.params.source_mode, 0)
    if w.params.source_mode == GEL_SRC_HYBRID:
        if s == 4 and not moving:
            src = 4 + 24 * s
        else:
            src = 4 + fallback_fraction * 19 * s + (1.0 - fallback_fraction) * 5 * s
    return 2 * nv * s + geom + src


def run_gpu(w, K, W, label):
    from gelios_2d_amd.setka import Setka
    nc = w.grid.ncells
    s = Setka(w.grid, w.params)
    if getattr(w, "motion", None):
        s.set_motion(*w.motion)
    s.upload(w.prim0)
    if getattr(w, "src", None) is not None:
        s.set_sources(w.src)
    if getattr(w, "tracking", None) is not None:
        s.set_tracking(w.tracking, w.rails)
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


def run_ensemble(steps, refresh_every, log_prefix="ensemble"):
    """Build the 64 cases, fill their source tables from the network (one fused launch
    for all cases, every cell judged), autotune, then time `steps` steps: HIP-graph
    replays in which the sources of every case are refreshed from the network, with
    the analytic judge, after every `refresh_every` steps; cells the judge has sent to
    the analytic closure keep it (judge 2), so that a refresh evaluates the network on
    the others only.  Returns (result dict, final states, specs, steps run before the
    timed region)."""
    from gelios_2d_amd.batch import Batch
    from gelios_2d_amd.configs import ensemble_cases
    from gelios_2d_amd.setka import Setka
    from gelios_2d_amd.surrogate import SourceModel
    from gelios_2d_amd.train_surrogate import DEFAULT_OUT
    import math
    import numpy as np
    model = SourceModel.load(DEFAULT_OUT)
    specs = ensemble_cases(64, seed=5, steps=steps)
    cases = []
    for w, f in specs:
        s = Setka(w.grid, w.params)
        s.upload(w.prim0)
        s.set_sources(f)
        s.set_source_table(np.zeros((5, s.ncells)))
        cases.append(s)
    ncells = sum(s.ncells for s in cases)
    b = Batch(cases)
    nfall = b.refresh(model, tol=1e-3, judge=True)
    rows, wg, ms_trial = b.autotune(12)
    b.set_source_model(model, refresh_every, tol=1e-3, judge=2)
    cyc = 6 * refresh_every // math.gcd(6, refresh_every)
    log(f"bench[{log_prefix}]: {len(cases)} cases, {ncells} cells; autotuner rows {rows} workgroups {wg} "
        f"({ms_trial * 1e3:.1f} us/step sweep only); initial judge fallback {nfall}/{ncells} cells; "
        f"source refresh every {refresh_every} steps, graph cycle {cyc} steps")
    b.step(cyc, True)                             # warm-up (graph capture)
    b.source_stats()
    t0 = time.perf_counter()
    dev_ms = b.step(steps, True)
    wall = time.perf_counter() - t0
    nref, nfb = b.source_stats()
    finals = [s.download() for s in cases]
    # device time of the sweep kernel alone: HIP events on the batch stream around
    # single-step launches (a graph replay would add its launch gaps)
    b.set_source_model(None, 0)
    sweep_ms = sum(b.step(1, False) for _ in range(60)) / 60
    # device time of the refresh alone, as the timed steps run it: one fused launch
    # over all cases
    t1 = time.perf_counter()
    reps = 10
    for _ in range(reps):
        b.refresh(model, tol=1e-3, judge=2)
    refresh_ms = (time.perf_counter() - t1) * 1e3 / reps
    return ({"value": ncells * steps / wall, "wall_s": wall, "cells": ncells, "cases": len(cases),
             "autotune": {"rows": rows, "workgroups": wg, "sweep_us_per_step": ms_trial * 1e3},
             "fallback_cells_initial": nfall, "refreshes": nref,
             "fallback_fraction": nfb / max(1, nref * ncells), "refresh_ms_host_timed": refresh_ms,
             "kernel_ms_per_step": dev_ms / steps, "sweep_ms_per_step": sweep_ms}, finals, specs, cyc)


def run_ensemble_analytic(specs, steps, cycle):
    """The same cases in FP32 bulk with the analytic closure at every step and no
    network (GEL_SRC_MK1), batched and graph-replayed like the timed run: what the
    learned source is measured against."""
    from gelios_2d_amd.batch import Batch
    from gelios_2d_amd.params import GEL_SRC_MK1
    from gelios_2d_amd.setka import Setka
    cases = []
    for w, f in specs:
        prm = w.params.copy()
        prm.source_mode = GEL_SRC_MK1
        s = Setka(w.grid, prm)
        s.upload(w.prim0)
        s.set_sources(f)
        cases.append(s)
    b = Batch(cases)
    b.autotune(12)
    b.step(cycle, True)
    t0 = time.perf_counter()
    b.step(steps, True)
    wall = time.perf_counter() - t0
    return {"value": sum(s.ncells for s in cases) * steps / wall, "ms_per_step": wall * 1e3 / steps}


def run_ensemble_reference(specs, steps):
    """The same cases in FP64 with the analytic sources every step (batched)."""
    from gelios_2d_amd.batch import Batch
    from gelios_2d_amd.params import GEL_PREC_FP64, GEL_SRC_MK1
    from gelios_2d_amd.setka import Setka
    cases = []
    for w, f in specs:
        prm = w.params.copy()
        prm.precision = GEL_PREC_FP64
        prm.source_mode = GEL_SRC_MK1
        s = Setka(w.grid, prm)
        s.upload(w.prim0)
        s.set_sources(f)
        cases.append(s)
    b = Batch(cases)
    b.step(steps, True)
    return [s.download() for s in cases]


def cpu_baseline(w, budget_s=12.0):
    """The CPU restatement of the reference sweep (per-cell gather over faces, every
    interior face solved from both sides, FP64, OpenMP over cells, as
    Setka.cpp:10148) on the host cores, over a bounded number of steps."""
    from oracle import oracle as ora
    threads = min(16, os.cpu_count() or 1)
    motion = getattr(w, "motion", None)
    if getattr(w, "tracking", None) is not None:
        tr = ora.TrackedOracle(w.grid, w.params, w.tracking, w.rails)
        sw = tr.sw
    else:
        tr = None
        sw = ora.OracleSweep.from_polar(w.grid, w.params)
        if getattr(w, "src", None) is not None:
            sw.src = w.src
    x0, y0 = w.grid.x, w.grid.y
    amp, period, phase = motion if motion else (0.0, 1, 0.0)

    def one(cur, n):
        if tr is not None:
            x2, y2, st = ora.track_nodes(tr.t, w.grid.ni, w.grid.nj, tr.x, tr.y, cur, tr.rails, w.params.gamma, sw.dt)
            sw.set_nodes(tr.x, tr.y, x2, y2)
            nxt, _ = sw.step(cur, nthreads=threads)
            tr.x, tr.y = x2, y2
            return nxt
        if motion:
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
    out = {"value": w.grid.ncells * steps / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
           "sample": f"{steps} steps of {w.name} ({w.grid.ncells} cells), FP64, OpenMP {threads} threads"}
    # fairness: the compiled reference sweep against the port on a grid both run (the
    # reference has no MHD sweep), same threads; where the harness was built
    from oracle import time_reference as tr
    if tr.available():
        m = tr.measure(200, 100, 20, threads)
        out["fairness"] = {"workload": "reference gas HLL sweep, 200x100 polar grid with C_1 wind rows, 20 steps",
                           "reference_cell_updates_per_s": m["reference"], "port_cell_updates_per_s": m["port"],
                           "port_over_reference": m["port"] / m["reference"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="ensemble")
    ap.add_argument("--refresh-every", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    from gelios_2d_amd import configs
    from gelios_2d_amd.params import GEL_PREC_FP32_BULK, GEL_PREC_FP64
    from gelios_2d_amd.state import l1_rel, linf_rel

    wl = args.workload
    W = args.warmup
    extra = {}
    if wl == "ensemble":
        import numpy as np
        K = args.steps if args.steps is not None else 500
        w64 = configs.hlld_mhd_800x400(GEL_PREC_FP64)
        log(f"bench: batched ensemble, {K} steps; FP64 plain reference workload {w64.name}")
        r64 = run_gpu(w64, 1000, W, "hlld-800x400-fp64-plain")
        r5, finals, specs, warm = run_ensemble(K, args.refresh_every)
        # the timed run did `warm` warm-up steps first: the reference runs K + warm steps
        ref = run_ensemble_reference(specs, K + warm)
        num = {0: 0.0, 1: 0.0}
        den = {0: 0.0, 1: 0.0}
        emax = {0: 0.0, 1: 0.0}
        vmax = {0: 0.0, 1: 0.0}
        for a, b in zip(finals, ref):
            for k in (0, 1):
                num[k] += float(np.abs(a[k] - b[k]).sum())
                den[k] += float(np.abs(b[k]).sum())
                emax[k] = max(emax[k], float(np.abs(a[k] - b[k]).max()))
                vmax[k] = max(vmax[k], float(np.abs(b[k]).max()))
        l1 = {"density": num[0] / den[0], "pressure": num[1] / den[1]}
        linf = {"density": emax[0] / vmax[0], "pressure": emax[1] / vmax[1]}
        log(f"bench[ensemble]: {r5['value']:.4e} cell-updates/s ({r5['wall_s']:.3f} s incl. source refreshes); "
            f"vs FP64 plain hlld-800x400 {r64['value']:.4e} -> x{r5['value'] / r64['value']:.2f}; L1 vs FP64 analytic {l1}; "
            f"Linf {linf}; "
            f"judge fallback fraction {r5['fallback_fraction']:.3f}")
        ra = run_ensemble_analytic(specs, K, warm)
        log(f"bench[ensemble]: same cases, FP32 bulk with the analytic closure only: {ra['value']:.4e} cell-updates/s")
        B = bytes_per_cell_update(specs[0][0], r5["fallback_fraction"])
        achieved = B * r5["cells"] / (r5["sweep_ms_per_step"] * 1e-3) / 1e9
        main_r = {"value": r5["value"], "wall_s": r5["wall_s"],
                  "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": achieved / HBM_PEAK_GBS, "traffic": ENSEMBLE_TRAFFIC_BYTES,
                               "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE)",
                               "bytes_per_cell_update": B, "kernel_avg_us": r5["sweep_ms_per_step"] * 1e3}}
        main_w = None
        extra = {
            "fp64_plain": {"workload": w64.name, "value": r64["value"], "ms_per_step": r64["wall_s"] * 1e3 / 1000,
                           "roofline": r64["roofline"]},
            "speedup_vs_fp64_plain": r5["value"] / r64["value"],
            "fp32_analytic_source_only": dict(ra, note="same cases, FP32 bulk, GEL_SRC_MK1 every step, no network"),
            "l1_vs_fp64": dict(l1, reference="same 64 cases, FP64, analytic sources every step", steps=K + warm),
            "linf_vs_fp64": dict(linf, definition="max |err| / max |FP64 value| over all cells of all cases"),
            "batch": {"cases": r5["cases"], "cells": r5["cells"], "autotune": r5["autotune"],
                      "source_refresh_every": args.refresh_every, "judge_tol": 1e-3,
                      "judge_fallback_cells_initial": r5["fallback_cells_initial"],
                      "refreshes_timed": r5["refreshes"], "judge_fallback_fraction": r5["fallback_fraction"],
                      "refresh_ms": r5["refresh_ms_host_timed"],
                      "kernel_ms_per_step_incl_refresh": r5["kernel_ms_per_step"]},
        }
        precision, dtype = "fp32-bulk + learned source (BF16 MFMA), batched, HIP graph", "f32"
        config = {"workload": "ensemble-64-cases-100x50-to-800x400-hlld-fp32-learned-source", "cells": r5["cells"], "cases": r5["cases"],
                  "physics": "mhd-hlld", "moving_grid": False, "precision": precision, "cfl": 0.4}
    elif wl in ("hlld-800x400", "hydrogen-1600x800"):
        if wl == "hlld-800x400":
            w64, w32 = configs.hlld_mhd_800x400(GEL_PREC_FP64), configs.hlld_mhd_800x400(GEL_PREC_FP32_BULK)
        else:
            w64 = configs.mhd_hydrogen_1600x800(GEL_PREC_FP64)
            w32 = configs.mhd_hydrogen_1600x800(GEL_PREC_FP32_BULK)
        K = args.steps if args.steps is not None else w64.steps
        log(f"bench: {w64.name} / {w32.name}, {w64.grid.ncells} cells, warmup {W}, steps {K}")
        r64 = run_gpu(w64, K, W, "fp64-plain")
        r32 = run_gpu(w32, K, W, "fp32-bulk")
        main_w, main_r = w32, r32
        extra = {
            "fp64_plain": {"value": r64["value"], "ms_per_step": r64["wall_s"] * 1e3 / K,
                           "roofline": r64["roofline"]},
            "speedup_fp32_vs_fp64": r32["value"] / r64["value"],
            "linf_vs_fp64": {"density": linf_rel(r32["final"], r64["final"], 0),
                             "pressure": linf_rel(r32["final"], r64["final"], 1),
                             "definition": "max |err| / max |FP64 value| over all cells"},
            "l1_vs_fp64": {"density": l1_rel(r32["final"], r64["final"], 0),
                           "pressure": l1_rel(r32["final"], r64["final"], 1),
                           "after_steps": W + K},
        }
        precision, dtype = "fp32-bulk (FP32 state and face arithmetic; FP64 geometry, accumulation, update)", "f32"
        config = None
    else:
        if wl == "hllc-200x100":
            w = configs.hllc_gas_200x100()
        elif wl == "large-mhd":
            w = configs.mhd_heliosphere(4096, 2048, 20, seed=2, name="hlld-fp64-4096x2048-moving")
        elif wl == "large-gas":
            w = configs.large()
        elif wl == "plasma-4h-800x400":
            w = configs.multifluid_heliosphere(800, 400, 200, seed=9, name="plasma-4h-hllc-fp64-800x400")
        elif wl == "tracked-800x400":
            w = configs.tracked_heliosphere(400, (200, 150, 150, 300), 200, seed=8, name="tracked-hll-fp64-799x400")
        else:
            raise SystemExit(f"unknown workload {wl}")
        K = args.steps if args.steps is not None else w.steps
        log(f"bench: {w.name}, {w.grid.ncells} cells, warmup {W}, steps {K}")
        main_w, main_r = w, run_gpu(w, K, W, "fp64-plain")
        precision, dtype = "fp64-plain", "f64"
        config = None

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
        "config": config if config is not None else {
            "workload": main_w.name, "cells": main_w.grid.ncells, "grid": [main_w.grid.ni, main_w.grid.nj],
            "physics": {1: "mhd-hlld", 2: "plasma+4H-hllc"}.get(main_w.params.physics,
                                                                 "gas-hllc" if main_w.params.flux == 1 else "gas-hll"),
            "moving_grid": bool(getattr(main_w, "motion", None) or getattr(main_w, "tracking", None)),
            "fitted_surfaces": getattr(main_w, "tracking", None) is not None,
            "precision": precision, "cfl": main_w.params.cfl},
        "roofline": main_r["roofline"],
    }
    line.update(extra)
    if not args.no_cpu:
        cw = main_w if wl not in ("hlld-800x400", "hydrogen-1600x800", "ensemble") else w64
        line["cpu_baseline"] = cpu_baseline(cw)
        log(f"bench: cpu baseline {line['cpu_baseline']['value']:.3e} cell-updates/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
