"""Train the learned charge-exchange source on the GPU and evaluate it.

    python -m gelios_2d_amd.train_surrogate [--steps 30000] [--out PATH] [--targets mc|closure]

Training data: 16 seeded inflow parameter sets (Mach 1.5-3, n_H 0.05-0.2) on 500 x 250
cells each (2M samples).  Targets: with --targets mc (default) the hydrogen Monte Carlo
on the device (gel_mc_run, 2e7 primaries per set): its moments are the hydrogen
features, its k factors (its momentum and energy sources over the closure's) the
targets, and the network gives factors on the closure (GEL_MLP_OUT_FACTOR); with
--targets closure the closure's sources over seeded hydrogen populations.
Evaluation on two held-out inflow sets: the network and the closure (Cell::Get_Sourse_MK1
without the Monte Carlo's k factors) against the Monte Carlo's own sources, volume
weighted, and the share of cells the device judge accepts.
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import surrogate as sg
from .params import GEL_SRC_TABLE
from .setka import Setka

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "source_mlp_16_inflow_sets.npz")


def per_cell_error(model: sg.SourceModel, setka: Setka, fields) -> np.ndarray:
    """max over components of |q_net - q_analytic| / max(|q_analytic|, sd * scale)
    -- the judge's measure (csrc/surrogate.hip, finalize_kernel)."""
    x = sg.features(setka)
    q = sg.analytic_sources(setka)
    scale = sg.target_scale(setka, fields)
    y = model.mlp.forward((x - model.mu) / model.sd)
    qn = (y * model.ysd + model.ymu).T * scale
    den = np.maximum(np.abs(q), model.ysd[:, None] * scale)
    return np.where(den > 0, np.abs(qn - q) / np.where(den > 0, den, 1.0), 0.0).max(axis=0)


def cell_volumes(g) -> np.ndarray:
    """r-weighted cell volumes of a polar grid (Pappus)."""
    x, y = g.x.reshape(g.ni + 1, g.nj + 1), g.y.reshape(g.ni + 1, g.nj + 1)
    px = np.stack([x[:-1, :-1], x[1:, :-1], x[1:, 1:], x[:-1, 1:]])
    py = np.stack([y[:-1, :-1], y[1:, :-1], y[1:, 1:], y[:-1, 1:]])
    G = sum((py[k] + py[(k + 1) % 4]) * (px[k] * py[(k + 1) % 4] - py[k] * px[(k + 1) % 4]) for k in range(4))
    return (np.abs(G) / 6.0).ravel()


def evaluate_mc(model: sg.SourceModel, mach, nh, seed, n_primary):
    """Network and closure against the Monte Carlo on a held-out inflow set: volume-weighted
    relative L1 of each momentum / energy source, with two Monte Carlo runs of different
    seeds giving the noise floor."""
    w, f = sg.workload_with_hydrogen(500, 250, 0, mach, nh, seed, moving=False)
    e = Setka(w.grid, w.params)           # 300 steps with the closure over the seeded populations
    e.upload(w.prim0)
    e.set_sources(f)
    e.Go_5_komponent__MK2(300)
    prm = w.params.copy()
    prm.source_mode = GEL_SRC_TABLE       # the judge fills the table
    s = Setka(w.grid, prm)
    s.upload(e.download())
    e.close()
    f2, q_mc2, _ = sg.mc_sources(s, w, nh, n_primary, seed + 17)
    f1, q_mc, st = sg.mc_sources(s, w, nh, n_primary, seed)
    vol = cell_volumes(w.grid)
    q_net = model.predict(s, f1)
    q_clo = sg.closure_k1(s, f1)

    def l1(a, b, m):
        return float((np.abs(a[m] - b[m]) * vol).sum() / ((np.abs(b[m]) * vol).sum() + 1e-300))
    rows = {}
    for name, q in (("network", q_net), ("closure_k1", q_clo), ("mc_other_seed", q_mc2)):
        rows[name] = [l1(q, q_mc, m) for m in (1, 2, 3)]
    # the judge's measure against the closure with the Monte Carlo's k factors
    qa = sg.analytic_sources(s)
    den = np.abs(qa[1:4])
    err = np.where(den > 0, np.abs(q_net[1:4] - qa[1:4]) / np.where(den > 0, den, 1.0), 0.0).max(axis=0)
    rows["judge_pass"] = {t: float(np.mean(err <= t)) for t in (1e-3, 1e-2, 3e-2, 1e-1)}
    return rows, st, s, f1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--targets", default="mc", choices=("mc", "closure"))
    ap.add_argument("--primaries", type=int, default=20_000_000)
    args = ap.parse_args()
    t0 = time.time()
    mode = sg.OUT_FACTOR if args.targets == "mc" else sg.OUT_SOURCE
    if args.targets == "mc":
        X, Y = sg.training_data_mc(ni=500, nj=250, n_sets=16, seed=4, max_evolve=600, n_primary=args.primaries,
                                   mode=mode)
    else:
        X, Y = sg.training_data(ni=500, nj=250, n_sets=16, seed=4, max_evolve=600)
    print(f"training set {X.shape[0]} samples x {X.shape[1]} features ({time.time() - t0:.1f} s)", flush=True)
    t0 = time.time()
    model = sg.train_source_model(X, Y, steps=args.steps, batch=4096, lr=1e-3, mode=mode)
    print(f"{args.steps} Adam steps, batch 4096, in {time.time() - t0:.1f} s; normalised MSE "
          f"{model.loss_history[:5].mean():.3e} -> {model.loss_history[-200:].mean():.3e}", flush=True)
    model.save(args.out)
    if args.targets == "mc":
        for mach, nh, seed in ((2.2, 0.12, 991), (1.7, 0.07, 992)):
            rows, st, s, f = evaluate_mc(model, mach, nh, seed, args.primaries)
            print(f"held-out Mach {mach} n_H {nh}: Monte Carlo {st['atoms']} atoms, {st['segments']} segments, "
                  f"{st['ms']:.0f} ms", flush=True)
            for name in ("network", "closure_k1", "mc_other_seed"):
                print(f"  volume-weighted relative L1 vs the Monte Carlo (mom x, mom y, energy) {name}: "
                      + " ".join(f"{e:.3e}" for e in rows[name]), flush=True)
            print("  cells within the judge's tolerance (network vs closure x Monte Carlo k): "
                  + " ".join(f"{v:.3f}@{t:g}" for t, v in rows["judge_pass"].items()), flush=True)
            for tol in (1e-3, 3e-2):
                print(f"  device judge at {tol:g}: {model.refresh(s, tol=tol)} of {s.ncells} cells fall back",
                      flush=True)
        return
    for mach, nh, seed in ((2.2, 0.12, 991), (1.7, 0.07, 992)):
        w, f = sg.workload_with_hydrogen(500, 250, 0, mach, nh, seed)
        prm = w.params.copy()
        prm.source_mode = GEL_SRC_TABLE
        s = Setka(w.grid, prm)
        s.upload(w.prim0)
        s.set_sources(f)
        e = per_cell_error(model, s, f)
        print(f"held-out Mach {mach} n_H {nh}: error median {np.median(e):.2e} p90 {np.quantile(e, 0.9):.2e} "
              f"p99 {np.quantile(e, 0.99):.2e}; accepted " +
              " ".join(f"{np.mean(e <= t):.3f}@{t:g}" for t in (1e-3, 3e-3, 1e-2, 3e-2)), flush=True)
        print(f"  device judge at 1e-3: {model.refresh(s, tol=1e-3)} of {s.ncells} cells fall back", flush=True)


if __name__ == "__main__":
    main()
