"""Benchmark of the Godunov flux-and-update sweep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one pass of the sweep over the whole grid (one launch of the sweep
kernel); a cell-update is one cell advanced by one step.  Inputs are resident in
HBM when the timed region starts.  Prints one JSON line.

Workloads
  ensemble (default)  the metric's comparison: 64 heliosphere cases (100x50 .. 800x400
                   cells, seeded inflow / field / hydrogen), FP32-bulk HLLD with the
                   learned charge-exchange source, one persistent launch per step,
                   HIP-graph replay, launch shape from the autotuner (value); beside it
                   the FP64 plain sweep of the same cases (analytic source every step,
                   batched, graph-replayed: the ratio's denominator), the same cases in
                   FP32 without the network, the density / pressure L1 and Linf errors
                   after the workload's 500 steps against the FP64 run, and the FP64
                   plain sweep of hlld-800x400
  learned-800x400  the learned-source form of hlld-800x400: four frozen hydrogen
                   populations, the network's charge-exchange source refreshed every 10
                   steps (hybrid), FP64 state, 1000 steps, beside FP64 with the analytic
                   closure every step
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
  mesh-800x400     the gas heliosphere (HLLC, twenty rows of C_1 wind cells, FP64, 200
                   steps) given as a general mesh: the general-mesh kernels (face table,
                   per-face and per-cell launches); structured-800x400 runs the same on
                   the structured marching kernel; mesh-800x400-fp32 the mesh in FP32 bulk
  block-731        the reference's own block grid (Setka(6,3,4,8,24,6,8,8), 731 cells, read
                   from its checkpoint in tests/golden/sweep_block.npz), HLL, the
                   reference's constants, 200 steps
  block-731-tracked  the same grid as the reference's constructor builds it
                   (tests/golden/sweep_block_tracked.npz) with its fitted surfaces moving:
                   every step Move_surface over the three lines, Move_Setka_Calculate_2
                   over the A, B, C, D rails, the sweep; 200 steps
  mc-200x100       the hydrogen Monte Carlo (gel_mc_run) on the hllc-200x100 heliosphere's
                   plasma: a step is one run of 4e6 primary atoms with their secondaries;
                   the line's value is cell segments flown per second (its own metric)
  mc-block-731     the same on the reference's own block grid and its checkpoint's plasma
                   (the atoms walking each cell's face slots)
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
# plus WRITE_SIZE, in KiB per dispatch, mean over the timed graph-replayed steps
# (profiles/sweep_batch_ensemble_64_cases_fp32_hybrid.md)
ENSEMBLE_TRAFFIC_BYTES = (2 * 556549 + 200796) * 1024
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
    moving = bool(getattr(w, "motion", None) or getattr(w, "tracking", None) or getattr(w, "mesh_rails", None))
    geom = 32 if moving else 16
    src = {GEL_SRC_TABLE: 5 * s, GEL_SRC_MK1: 19 * s}.get(w.params.source_mode, 0)
    if w.params.source_mode == GEL_SRC_HYBRID:
        if s == 4 and not moving:
            src = 4 + 24 * s
        else:
            src = 4 + fallback_fraction * 19 * s + (1.0 - fallback_fraction) * 5 * s
    return 2 * nv * s + geom + src


def dict_params(w, precision, source_mode):
    """A stand-in of workload w with another precision and source mode (for the byte
    count)."""
    import copy
    v = copy.copy(w)
    v.params = w.params.copy()
    v.params.precision = precision
    v.params.source_mode = source_mode
    return v


def run_gpu(w, K, W, label):
    """W warm-up steps, then K steps timed by the host clock around the call and by HIP
    events on the context's stream around the same call (timing mode 2: the per-step
    device time of back-to-back launches)."""
    from gelios_2d_amd.setka import Setka
    nc = w.grid.ncells
    s = Setka.from_mesh(w.mesh, w.params) if getattr(w, "mesh", None) is not None else Setka(w.grid, w.params)
    if getattr(w, "motion", None):
        s.set_motion(*w.motion)
    s.upload(w.prim0)
    if getattr(w, "src", None) is not None:
        s.set_sources(w.src)
    if getattr(w, "mesh_rails", None) is not None:
        s.set_mesh_tracking(w.mesh_tracking, w.mesh_rails)
    elif getattr(w, "tracking", None) is not None:
        s.set_tracking(w.tracking, w.rails)
    s.Go_5_komponent__MK2(W)
    s.sync()
    s.enable_timing(2)
    t0 = time.perf_counter()
    s.Go_5_komponent__MK2(K)
    st = s.sync()
    wall = time.perf_counter() - t0
    avg_ms, n = s.kernel_timing()
    s.enable_timing(0)
    value = nc * K / wall
    log(f"bench[{label}]: {K} steps in {wall:.4f} s -> {value:.4e} cell-updates/s, dt_next {st.dt_next:.3e}, "
        f"clamps rho {st.n_rho_clamp} p {st.n_p_clamp}")
    final = s.download()
    B = bytes_per_cell_update(w)
    achieved = B * nc / (avg_ms * 1e-3) / 1e9
    log(f"bench[{label}]: device time per step {avg_ms * 1e3:.2f} us over the {n} timed steps, "
        f"{achieved:.1f} GB/s algorithmic ({B} B/cell-update)")
    s.close()
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_cell_update": B,
            "kernel_avg_us": avg_ms * 1e3,
            "kernel_time": "HIP events around the timed call on the context's stream / steps"}
    return {"value": value, "wall_s": wall, "final": final, "roofline": roof, "steps_done": K + W,
            "dt_next": st.dt_next, "n_p_clamp": st.n_p_clamp, "n_rho_clamp": st.n_rho_clamp}


def run_learned(w, K, W, every, tol=1e-3):
    """A single context with the learned source (hybrid): the network refreshes the
    sources after every `every`-th step (sticky judge), host-driven between the sweep
    calls; W warm-up steps, then K steps timed (host clock; HIP events around each
    sweep call, refreshes excluded, for the kernel's per-step time)."""
    import numpy as np
    from gelios_2d_amd.setka import Setka
    from gelios_2d_amd.surrogate import SourceModel
    from gelios_2d_amd.train_surrogate import DEFAULT_OUT
    model = SourceModel.load(DEFAULT_OUT)
    s = Setka(w.grid, w.params)
    s.set_motion(*w.motion)
    s.upload(w.prim0)
    s.set_sources(w.src)
    s.set_source_table(np.zeros((5, s.ncells)))
    nfall0 = model.refresh(s, tol, judge=True)
    model.advance(s, W, every, tol)
    s.sync()
    s.enable_timing(2)
    t0 = time.perf_counter()
    nfall = model.advance(s, K, every, tol)
    st = s.sync()
    wall = time.perf_counter() - t0
    avg_ms, n = s.kernel_timing()
    s.enable_timing(0)
    final = s.download()
    s.close()
    B = bytes_per_cell_update(w, (nfall or 0) / w.grid.ncells)
    achieved = B * w.grid.ncells / (avg_ms * 1e-3) / 1e9
    log(f"bench[learned]: {K} steps in {wall:.4f} s -> {w.grid.ncells * K / wall:.4e} cell-updates/s; sweep "
        f"{avg_ms * 1e3:.2f} us/step; flagged {nfall0} -> {nfall} of {w.grid.ncells}")
    return {"value": w.grid.ncells * K / wall, "wall_s": wall, "final": final, "n_p_clamp": st.n_p_clamp,
            "fallback_initial": nfall0, "fallback_last": nfall,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_cell_update": B,
                         "kernel_avg_us": avg_ms * 1e3,
                         "kernel_time": "HIP events around the sweep calls of the timed region / steps"}}


def batch_cases(specs, precision=None, source_mode=None):
    """One context per ensemble case (uploaded state, hydrogen moments, an empty source
    table where the mode reads one)."""
    import numpy as np
    from gelios_2d_amd.params import GEL_SRC_HYBRID, GEL_SRC_TABLE
    from gelios_2d_amd.setka import Setka
    cases = []
    for w, f in specs:
        prm = w.params.copy()
        if precision is not None:
            prm.precision = precision
        if source_mode is not None:
            prm.source_mode = source_mode
        s = Setka(w.grid, prm)
        s.upload(w.prim0)
        s.set_sources(f)
        if prm.source_mode in (GEL_SRC_TABLE, GEL_SRC_HYBRID):
            s.set_source_table(np.zeros((5, s.ncells)))
        cases.append(s)
    return cases


def timed_batch(b, K, W):
    """W untimed warm-up steps, the graphs of a K-step call captured (nothing run),
    then exactly K steps timed: host clock around the call, HIP events on the batch
    stream inside it.  Returns (wall seconds, device milliseconds)."""
    if W:
        b.step(W, True)
    b.source_stats()
    b.prepare(K)
    t0 = time.perf_counter()
    dev_ms = b.step(K, True)
    return time.perf_counter() - t0, dev_ms


def sweep_us_per_step(b, n=60):
    """Device time per step of the batched sweep kernel alone: HIP events around one
    call of n graph-replayed steps with no source refresh (kernel back to back)."""
    b.set_source_model(None, 0)
    b.prepare(n)
    return b.step(n, True) * 1e3 / n


def run_ensemble(specs, K, W, refresh_every, tol=1e-3):
    """FP32 bulk + the learned source (hybrid: the network's table, the analytic closure
    in the cells its judge rejected), batched, graph-replayed; the network refreshes
    every case's sources after every `refresh_every` steps inside the graphs."""
    from gelios_2d_amd.batch import Batch
    from gelios_2d_amd.surrogate import SourceModel
    from gelios_2d_amd.train_surrogate import DEFAULT_OUT
    model = SourceModel.load(DEFAULT_OUT)
    cases = batch_cases(specs)
    ncells = sum(s.ncells for s in cases)
    b = Batch(cases)
    nfall = b.refresh(model, tol=tol, judge=True)
    rows, wg, ms_trial = b.autotune(12)
    b.set_source_model(model, refresh_every, tol=tol, judge=2)
    wall, dev_ms = timed_batch(b, K, W)
    nref, nfb = b.source_stats()
    log(f"bench[ensemble]: {len(cases)} cases, {ncells} cells; autotuner rows {rows} workgroups {wg}; "
        f"initial judge fallback {nfall}/{ncells}; {K} steps in {wall:.4f} s ({nref} refreshes)")
    sweep_us = sweep_us_per_step(b)
    t1 = time.perf_counter()
    for _ in range(5):
        b.refresh(model, tol=tol, judge=2)
    refresh_ms = (time.perf_counter() - t1) * 1e3 / 5
    for s in cases:
        s.close()
    return {"value": ncells * K / wall, "wall_s": wall, "cells": ncells, "cases": len(cases),
            "autotune": {"rows": rows, "workgroups": wg, "sweep_us_per_step": ms_trial * 1e3},
            "fallback_cells_initial": nfall, "refreshes": nref,
            "fallback_fraction": nfb / max(1, nref * ncells), "refresh_ms_host_timed": refresh_ms,
            "device_ms_per_step": dev_ms / K, "sweep_us_per_step": sweep_us}


def run_ensemble_plain(specs, K, W, precision, source_mode, label):
    """The same cases without the network: `source_mode` (the analytic closure every
    step) in `precision`, batched, autotuned and graph-replayed like the learned run."""
    from gelios_2d_amd.batch import Batch
    cases = batch_cases(specs, precision, source_mode)
    ncells = sum(s.ncells for s in cases)
    b = Batch(cases)
    rows, wg, _ = b.autotune(12)
    wall, dev_ms = timed_batch(b, K, W)
    sweep_us = sweep_us_per_step(b)
    log(f"bench[{label}]: {K} steps in {wall:.4f} s -> {ncells * K / wall:.4e} cell-updates/s "
        f"(rows {rows}, workgroups {wg}; sweep {sweep_us:.1f} us/step)")
    for s in cases:
        s.close()
    return {"value": ncells * K / wall, "ms_per_step": wall * 1e3 / K, "sweep_us_per_step": sweep_us,
            "cells": ncells, "autotune": {"rows": rows, "workgroups": wg}}


def ensemble_accuracy(specs, steps, refresh_every, tol=1e-3):
    """Density / pressure error of the learned FP32 run against FP64 with the analytic
    sources every step, both from the initial states over `steps` steps (the workload's
    length), untimed."""
    import numpy as np
    from gelios_2d_amd.batch import Batch
    from gelios_2d_amd.params import GEL_PREC_FP64, GEL_SRC_MK1
    from gelios_2d_amd.surrogate import SourceModel
    from gelios_2d_amd.train_surrogate import DEFAULT_OUT
    model = SourceModel.load(DEFAULT_OUT)
    a = batch_cases(specs)
    A = Batch(a)
    A.refresh(model, tol=tol, judge=True)
    A.set_source_model(model, refresh_every, tol=tol, judge=2)
    A.step(steps, True)
    nref, nfb = A.source_stats()
    finals = [s.download() for s in a]
    for s in a:
        s.close()
    r = batch_cases(specs, GEL_PREC_FP64, GEL_SRC_MK1)
    Bt = Batch(r)
    Bt.step(steps, True)
    ref = [s.download() for s in r]
    for s in r:
        s.close()
    num, den, emax, vmax = [0.0, 0.0], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0]
    for x, y in zip(finals, ref):
        for k in (0, 1):
            num[k] += float(np.abs(x[k] - y[k]).sum())
            den[k] += float(np.abs(y[k]).sum())
            emax[k] = max(emax[k], float(np.abs(x[k] - y[k]).max()))
            vmax[k] = max(vmax[k], float(np.abs(y[k]).max()))
    ncells = sum(int(np.asarray(w.prim0).shape[1]) for w, _ in specs)
    return ({"density": num[0] / den[0], "pressure": num[1] / den[1], "steps": steps,
             "reference": "same 64 cases, FP64, analytic sources every step"},
            {"density": emax[0] / vmax[0], "pressure": emax[1] / vmax[1],
             "definition": "max |err| / max |FP64 value| over all cells of all cases", "steps": steps},
            nfb / max(1, nref * ncells))


def cpu_threads():
    """Threads the host grants this job: OMP_NUM_THREADS where set (a one-GPU job's CPU
    share), else every CPU the process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def host_cpu():
    """lscpu's model, sockets, cores per socket and threads per core."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(specs, budget_s=12.0):
    """The CPU restatement of the reference sweep (per-cell gather over faces, every
    interior face solved from both sides, FP64, OpenMP over cells, as Setka.cpp:10148)
    on the host cores, over a bounded sample of the ensemble: its cases in order, FP64
    with the analytic charge-exchange source every step (the plain form of the
    workload), a few steps each, until the time budget is spent."""
    import numpy as np
    from oracle import oracle as ora
    from gelios_2d_amd.params import GEL_PREC_FP64, GEL_SRC_MK1
    threads = cpu_threads()
    cells = 0
    t_total = 0.0
    passes = 0
    sweeps = []
    for w, f in specs:
        prm = w.params.copy()
        prm.precision = GEL_PREC_FP64
        prm.source_mode = GEL_SRC_MK1
        sw = ora.OracleSweep.from_polar(w.grid, prm)
        sw.src = np.ascontiguousarray(f, np.float64)
        cur, _ = sw.step(w.prim0, nthreads=threads)          # warm-up (threads, pages)
        sweeps.append([sw, cur, w.grid.ncells])
    while t_total < budget_s:                             # passes of 5 steps over every case
        for e in sweeps:
            t0 = time.perf_counter()
            for _ in range(5):
                e[1], _ = e[0].step(e[1], nthreads=threads)
            t_total += time.perf_counter() - t0
            cells += e[2] * 5
        passes += 1
    value = cells / t_total
    out = {"value": value, "unit": "cell-updates/s", "cores": threads, "kind": "port",
           "sample": f"all {len(specs)} ensemble cases, {passes} passes of 5 steps each ({5 * passes} steps per case), "
                     f"FP64 HLLD with the analytic charge-exchange source, OpenMP {threads} threads, "
                     f"{t_total:.1f} s timed",
           "host": host_cpu()}
    # the compiled reference against the port on a workload both run (the reference has
    # no MHD sweep), measured where the reference's sources are (oracle/time_reference.py)
    fj = os.path.join(ROOT, "oracle", "fairness.json")
    if os.path.exists(fj):
        with open(fj) as f:
            fair = json.load(f)["entries"]
        e = min(fair, key=lambda x: abs(x["threads"] - threads))
        out["fairness"] = {"port_over_reference": e["port_over_reference"], "threads": e["threads"],
                           "measured_on": e["cpu"], "workload": e["workload"],
                           "note": "measured in the build container, where the reference's sources are; "
                                   "the reference is not run on this host"}
        out["reference_equivalent_value"] = value / e["port_over_reference"]
    return out


def cpu_baseline_workload(w, budget_s=12.0):
    """The CPU restatement of the reference sweep (per-cell gather over faces, every
    interior face solved from both sides, FP64, OpenMP over cells, as
    Setka.cpp:10148) on the host cores, over a bounded number of steps."""
    import numpy as np
    from oracle import oracle as ora
    threads = cpu_threads()
    motion = getattr(w, "motion", None)
    mtr = None
    if getattr(w, "mesh_rails", None) is not None:
        from oracle import mesh_tracking as omt
        tr = None
        mtr = omt.MeshTrackedOracle(w.mesh, w.mesh_rails, w.mesh_tracking, w.params)
        sw = mtr.sw
    elif getattr(w, "tracking", None) is not None:
        tr = ora.TrackedOracle(w.grid, w.params, w.tracking, w.rails)
        sw = tr.sw
    elif getattr(w, "mesh", None) is not None:
        tr = None
        sw = ora.OracleSweep(w.mesh.general(), w.mesh.x, w.mesh.y, w.params)
        sw.ptype = np.ascontiguousarray(w.mesh.ptype, np.int32)
    else:
        tr = None
        sw = ora.OracleSweep.from_polar(w.grid, w.params)
        if getattr(w, "src", None) is not None:
            sw.src = w.src
    x0, y0 = w.grid.x, w.grid.y
    amp, period, phase = motion if motion else (0.0, 1, 0.0)

    def one(cur, n):
        if mtr is not None:
            x2, y2 = omt.move_nodes(mtr.mesh, mtr.rails, mtr.t, mtr.x, mtr.y, cur, sw.dt, w.params.gamma)
            sw.set_nodes(mtr.x, mtr.y, x2, y2)
            nxt, _ = sw.step(cur, nthreads=threads)
            mtr.x, mtr.y = x2, y2
            return nxt
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
    out["host"] = host_cpu()
    return out


# Algorithmic bytes per cell segment of the Monte Carlo: the cell's plasma (rho, cp, u, v,
# p: 40 B) and its four nodes (64 B) read, four moment tallies and three source tallies
# (read and written, 16 B each: 112 B); secondaries' queue traffic is not counted.
MC_BYTES_PER_SEGMENT = 40 + 64 + 112


def block_731_workload():
    """The reference's block grid Setka(6,3,4,8,24,6,8,8) and its state, read from its
    checkpoint (tests/golden/sweep_block.npz)."""
    from gelios_2d_amd import configs
    from gelios_2d_amd.checkpoint import read_alpha
    from gelios_2d_amd.mesh import mesh_from_alpha, state_from_alpha
    from gelios_2d_amd.params import reference_params
    import tempfile
    import numpy as np
    d = np.load(os.path.join(ROOT, "tests", "golden", "sweep_block.npz"))
    with tempfile.NamedTemporaryFile("wb", suffix=".txt") as f:
        f.write(d["b_ref_text"].tobytes())
        f.flush()
        alpha = read_alpha(f.name)
    m = mesh_from_alpha(alpha)
    w = configs.Workload("reference-block-grid-731-cells-hll-fp64", m, reference_params(), state_from_alpha(alpha),
                         200, None)
    w.mesh = m
    return w


def run_mc_workload(w, K, W, no_cpu, n_primary=4_000_000):
    """K timed runs of the Monte Carlo after W warm-up runs of a tenth of the atoms; on a
    general mesh when the workload carries one."""
    from gelios_2d_amd.mc import run_mc
    from gelios_2d_amd.params import mc_reference_params
    from gelios_2d_amd.setka import Setka
    mesh = getattr(w, "mesh", None)
    s = Setka.from_mesh(mesh, w.params) if mesh is not None else Setka(w.grid, w.params)
    s.upload(w.prim0)
    for k in range(max(W, 1)):
        run_mc(s, mc_reference_params(n_primary // 10, seed=100 + k))
    segs = atoms = lost = 0
    dev_ms = 0.0
    t0 = time.perf_counter()
    for k in range(K):
        _, st = run_mc(s, mc_reference_params(n_primary, seed=1 + k))
        segs += st["segments"]
        atoms += st["atoms"]
        lost += st["lost"]
        dev_ms += st["ms"]
        assert st["dropped"] == 0, st
    wall = time.perf_counter() - t0
    achieved = MC_BYTES_PER_SEGMENT * segs / (dev_ms * 1e-3) / 1e9
    line = {"metric": "cell segments/s of the hydrogen Monte Carlo on 1 MI355X", "value": segs / wall,
            "unit": "segments/s", "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": wall * 1e3 / K,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"mc-{w.name}", "cells": w.grid.ncells, "primaries_per_step": n_primary,
                       "atoms_per_step": atoms / K, "segments_per_step": segs / K,
                       "atoms_lost": lost},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "bytes_per_segment": MC_BYTES_PER_SEGMENT, "kernel_ms_per_step": dev_ms / K,
                         "note": "FP64 issue and divergence bound, not HBM: the fraction is what the "
                                 "algorithmic bytes would need"}}
    if not no_cpu:
        from oracle import oracle as ora
        import numpy as np
        n_cpu, elapsed, segs_cpu = 20000, 0.0, 0
        while elapsed < 8.0:
            t = time.perf_counter()
            pop = ora.mc_population(w.grid.cell_type, w.grid.ncells)
            if mesh is not None:
                _, st = ora.mc_run_mesh(mesh, w.prim0, pop, mc_reference_params(n_cpu, seed=7))
            else:
                _, st = ora.mc_run(w.grid, w.prim0, pop, mc_reference_params(n_cpu, seed=7))
            elapsed = time.perf_counter() - t
            segs_cpu = st["segments"]
            if elapsed < 8.0:
                n_cpu *= 2
        line["cpu_baseline"] = {"value": segs_cpu / elapsed, "unit": "segments/s", "cores": 1, "kind": "port",
                                "sample": f"{n_cpu} primaries of the same run, oracle/mc_oracle.c, one thread",
                                "host": host_cpu(),
                                "reference": "not run: the reference's MK_start_new needs about 8 h on 8 threads "
                                             "for its 411 x 50 primaries on its production grid"}
    return line


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
    from gelios_2d_amd.params import GEL_PREC_FP32_BULK, GEL_PREC_FP64, GEL_SRC_MK1
    from gelios_2d_amd.state import l1_rel, linf_rel

    wl = args.workload
    W = args.warmup
    extra = {}
    if wl == "ensemble":
        K = args.steps if args.steps is not None else 500
        specs = configs.ensemble_cases(64, seed=5, steps=500)
        log(f"bench: batched ensemble of {len(specs)} cases, warmup {W}, {K} timed steps")
        r5 = run_ensemble(specs, K, W, args.refresh_every)
        r64 = run_ensemble_plain(specs, K, W, GEL_PREC_FP64, GEL_SRC_MK1, "ensemble-fp64-plain")
        ra = run_ensemble_plain(specs, K, W, GEL_PREC_FP32_BULK, GEL_SRC_MK1, "ensemble-fp32-analytic")
        l1, linf, ffrac = ensemble_accuracy(specs, 500, args.refresh_every)
        w2 = configs.hlld_mhd_800x400(GEL_PREC_FP64)
        r2 = run_gpu(w2, 1000, W, "hlld-800x400-fp64-plain")
        log(f"bench[ensemble]: {r5['value']:.4e} cell-updates/s against FP64 plain on the same cases "
            f"{r64['value']:.4e} -> x{r5['value'] / r64['value']:.2f}; FP32 with the analytic closure only "
            f"{ra['value']:.4e}; L1 after 500 steps {l1}; Linf {linf}; fallback fraction {ffrac:.3f}")
        cells = r5["cells"]
        B = 136           # algorithmic bytes per cell-update of this workload (FP32 state, nodes, 8 H moments)
        B_in = bytes_per_cell_update(specs[0][0])   # what the kernel reads and writes per cell
        achieved = B * cells / (r5["sweep_us_per_step"] * 1e-6) / 1e9
        main_r = {"value": r5["value"], "wall_s": r5["wall_s"],
                  "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": achieved / HBM_PEAK_GBS, "traffic": ENSEMBLE_TRAFFIC_BYTES,
                               "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE)",
                               "bytes_per_cell_update": B, "kernel_avg_us": r5["sweep_us_per_step"],
                               "kernel_time": "HIP events on the batch stream around 60 graph-replayed "
                                              "steps of the sweep alone / 60",
                               "kernel_inputs_bytes_per_cell_update": B_in,
                               "kernel_inputs_frac": B_in * cells / (r5["sweep_us_per_step"] * 1e-6) / 1e9 / HBM_PEAK_GBS}}
        main_w = None
        B64 = bytes_per_cell_update(dict_params(specs[0][0], GEL_PREC_FP64, GEL_SRC_MK1))
        a64 = B64 * cells / (r64["sweep_us_per_step"] * 1e-6) / 1e9
        extra = {
            "fp64_plain": dict(r64, workload="the same 64 cases, FP64 state and arithmetic, analytic "
                                             "charge-exchange source every step, batched, HIP graph",
                               roofline={"bound": "hbm", "achieved": a64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                         "frac": a64 / HBM_PEAK_GBS, "bytes_per_cell_update": B64,
                                         "kernel_avg_us": r64["sweep_us_per_step"]}),
            "speedup_vs_fp64_plain": r5["value"] / r64["value"],
            "fp32_analytic_source_only": dict(ra, note="same cases, FP32 bulk, analytic closure every step, "
                                                       "no network"),
            "fp64_plain_800x400_moving": {"workload": w2.name, "value": r2["value"],
                                          "ms_per_step": r2["wall_s"] * 1e3 / 1000, "roofline": r2["roofline"],
                                          "p_clamps": r2["n_p_clamp"], "steps": 1000 + W},
            "l1_vs_fp64": l1,
            "linf_vs_fp64": linf,
            "batch": {"cases": r5["cases"], "cells": cells, "autotune": r5["autotune"],
                      "source_refresh_every": args.refresh_every, "judge_tol": 1e-3,
                      "judge_fallback_cells_initial": r5["fallback_cells_initial"],
                      "refreshes_timed": r5["refreshes"], "judge_fallback_fraction_timed": r5["fallback_fraction"],
                      "judge_fallback_fraction_500_steps": ffrac,
                      "refresh_ms": r5["refresh_ms_host_timed"],
                      "device_ms_per_step_incl_refresh": r5["device_ms_per_step"]},
        }
        precision, dtype = "fp32-bulk + learned source (BF16 MFMA, hybrid), batched, HIP graph", "f32"
        config = {"workload": "ensemble-64-cases-100x50-to-800x400-hlld-fp32-learned-source", "cells": cells,
                  "cases": r5["cases"], "physics": "mhd-hlld", "moving_grid": False, "precision": precision,
                  "cfl": 0.4}
    elif wl == "learned-800x400":
        w = configs.learned_mhd_800x400(GEL_PREC_FP64)
        w64 = configs.learned_mhd_800x400(GEL_PREC_FP64, GEL_SRC_MK1)
        K = args.steps if args.steps is not None else w.steps
        log(f"bench: {w.name} against {w64.name}, {w.grid.ncells} cells, warmup {W}, steps {K}")
        rl = run_learned(w, K, W, args.refresh_every)
        r64 = run_gpu(w64, K, W, "fp64-analytic")
        main_w, main_r = w, rl
        extra = {
            "fp64_plain": {"workload": w64.name, "value": r64["value"], "ms_per_step": r64["wall_s"] * 1e3 / K,
                           "roofline": r64["roofline"]},
            "speedup_vs_fp64_plain": rl["value"] / r64["value"],
            "l1_vs_fp64": {"density": l1_rel(rl["final"], r64["final"], 0),
                           "pressure": l1_rel(rl["final"], r64["final"], 1), "after_steps": W + K},
            "judge": {"flagged_initial": rl["fallback_initial"], "flagged_last_refresh": rl["fallback_last"],
                      "refresh_every": args.refresh_every, "tol": 1e-3},
        }
        precision, dtype = "fp64 state and arithmetic + learned source (BF16 MFMA, hybrid)", "f64"
        config = None
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
    elif wl == "mc-200x100":
        r = run_mc_workload(configs.hllc_gas_200x100(), args.steps if args.steps is not None else 3, W, args.no_cpu)
        print(json.dumps(r), flush=True)
        return
    elif wl == "mc-block-731":
        r = run_mc_workload(block_731_workload(), args.steps if args.steps is not None else 3, W, args.no_cpu)
        print(json.dumps(r), flush=True)
        return
    else:
        if wl == "hllc-200x100":
            w = configs.hllc_gas_200x100()
        elif wl == "large-mhd":
            w = configs.mhd_heliosphere(4096, 2048, 20, seed=2, name="hlld-fp64-4096x2048-moving")
        elif wl == "large-gas":
            w = configs.large()
        elif wl == "plasma-4h-800x400":
            w = configs.multifluid_heliosphere(800, 400, 200, seed=9, name="plasma-4h-hllc-fp64-800x400")
        elif wl in ("mesh-800x400", "mesh-800x400-fp32", "structured-800x400"):
            from gelios_2d_amd.mesh import mesh_from_polar
            from gelios_2d_amd.params import GEL_PREC_FP32_BULK
            fp32 = wl.endswith("fp32")
            w = configs.heliosphere(800, 400, 200, seed=1, flux=1,
                                    name=f"hllc-{'fp32-bulk' if fp32 else 'fp64'}-c1-800x400-{wl.split('-')[0]}")
            w.grid.cell_type[: 20 * w.grid.nj] = 0                  # twenty rows of C_1 wind cells
            if fp32:
                w.params.precision = GEL_PREC_FP32_BULK
            if wl.startswith("mesh"):
                w.mesh = mesh_from_polar(w.grid)
        elif wl == "block-731":
            w = block_731_workload()
        elif wl == "block-731-tracked":
            import numpy as np
            from gelios_2d_amd.mesh import MeshRails, mesh_from_slots
            from gelios_2d_amd.params import N_OUTER_SHOCK, reference_params, reference_tracking
            d = np.load(os.path.join(ROOT, "tests", "golden", "sweep_block_tracked.npz"))
            g = lambda k: d["b_" + k]   # noqa: E731
            m = mesh_from_slots(g("x"), g("y"), g("ptype"), g("ncorner"), g("corners"), g("cell_type"), g("nface"),
                                g("slots"))
            w = configs.Workload("reference-block-grid-731-cells-tracked-hll-fp64", m, reference_params(),
                                 g("prim0"), 200, None)
            w.mesh = m
            w.mesh_rails = MeshRails(g("rail_kind"), g("rail_m"), g("rail_s"), g("rail_offset"), g("rail_all"),
                                     g("rail_key"), (g("line_contact"), g("line_inner"), g("line_outer")))
            N = g("N")
            w.mesh_tracking = reference_tracking(outer_faces=N_OUTER_SHOCK + int(N[0]) + int(N[1]) + 1)
        elif wl == "tracked-800x400":
            w = configs.tracked_heliosphere(400, (200, 150, 150, 300), 200, seed=8, name="tracked-hll-fp64-799x400")
        else:
            raise SystemExit(f"unknown workload {wl}")
        K = args.steps if args.steps is not None else w.steps
        log(f"bench: {w.name}, {w.grid.ncells} cells, warmup {W}, steps {K}")
        fp32 = w.params.precision != 0
        main_w, main_r = w, run_gpu(w, K, W, "fp32-bulk" if fp32 else "fp64-plain")
        precision, dtype = ("fp32-bulk", "f32") if fp32 else ("fp64-plain", "f64")
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
            "workload": main_w.name, "cells": main_w.grid.ncells,
            "grid": [main_w.grid.ni, main_w.grid.nj] if hasattr(main_w.grid, "ni") else "general mesh",
            "physics": {1: "mhd-hlld", 2: "plasma+4H-hllc"}.get(main_w.params.physics,
                                                                 "gas-hllc" if main_w.params.flux == 1 else "gas-hll"),
            "moving_grid": bool(getattr(main_w, "motion", None) or getattr(main_w, "tracking", None)
                                or getattr(main_w, "mesh_rails", None)),
            "fitted_surfaces": (getattr(main_w, "tracking", None) is not None
                                or getattr(main_w, "mesh_rails", None) is not None),
            "precision": precision, "cfl": main_w.params.cfl},
        "roofline": main_r["roofline"],
    }
    line.update(extra)
    if not args.no_cpu:
        if wl == "ensemble":
            line["cpu_baseline"] = cpu_baseline(specs)
        else:
            line["cpu_baseline"] = cpu_baseline_workload(main_w if wl not in ("hlld-800x400", "hydrogen-1600x800",
                                                                               "learned-800x400") else w64)
        log(f"bench: cpu baseline {line['cpu_baseline']['value']:.3e} cell-updates/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
