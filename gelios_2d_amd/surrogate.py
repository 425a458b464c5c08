"""The learned charge-exchange source: a 14 -> 4 x 128 SiLU -> 5 MLP on MFMA.

Forward, backward and Adam run in libgelios.so (csrc/surrogate.hip): BF16 weights
and activations, FP32 accumulation.  This module is the host driver: parameter I/O,
training on the device, and the feature/target scaling of the source model.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

N_IN, N_OUT = 14, 5
_fp = ctypes.POINTER(ctypes.c_float)


def _f(a):
    return a.ctypes.data_as(_fp)


def _setup(L):
    """Signatures live in _lib.SIGNATURES; kept as the single place a handle is made."""
    return L


def _check(st):
    if st != 0:
        L = _lib.lib()
        raise _lib.GelError(f"{L.gel_strerror(st).decode()}: {L.gel_mlp_last_error().decode()}")


class MLP:
    """Device MLP with FP32 master parameters and BF16 compute copies."""

    def __init__(self, seed: int = 0):
        self._L = _setup(_lib.lib())
        self._h = ctypes.c_void_p()
        _check(self._L.gel_mlp_create(ctypes.byref(self._h), ctypes.c_uint64(seed)))
        self.n_params = self._L.gel_mlp_param_count()

    def close(self):
        if self._h:
            self._L.gel_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        _check(self._L.gel_mlp_get_params(self._h, _f(p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(self._L.gel_mlp_set_params(self._h, _f(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_IN)
        y = np.empty((x.shape[0], N_OUT), np.float32)
        _check(self._L.gel_mlp_forward(self._h, x.shape[0], _f(x), _f(y)))
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_IN)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, N_OUT)
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_float()
        _check(self._L.gel_mlp_grad(self._h, x.shape[0], _f(x), _f(t), _f(g), ctypes.byref(loss)))
        return loss.value, g

    def set_dataset(self, x, t) -> None:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_IN)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, N_OUT)
        _check(self._L.gel_mlp_set_dataset(self._h, x.shape[0], _f(x), _f(t)))

    def train(self, steps: int, batch: int, lr: float) -> np.ndarray:
        loss = np.empty(max(steps, 1), np.float32)
        _check(self._L.gel_mlp_train(self._h, int(steps), int(batch), float(lr), _f(loss)))
        return loss[:steps]


# ------------------------------------------------------------------ source model ---

def _setup_src(L):
    return L


def features(setka) -> np.ndarray:
    """Raw source-model features [ncells][14] of a Setka's current state."""
    L = _setup_src(_setup(_lib.lib()))
    x = np.empty((setka.ncells, N_IN), np.float32)
    _check(L.gel_surrogate_features(setka._ctx, _f(x)))
    return x


def analytic_sources(setka) -> np.ndarray:
    """Analytic charge-exchange sources [5][ncells] (mass, mom x, mom y, energy, tracer)."""
    L = _setup_src(_setup(_lib.lib()))
    q = np.empty((5, setka.ncells))
    _check(L.gel_source_analytic(setka._ctx, q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return q


def closure_k1(setka, fields) -> np.ndarray:
    """The closure [5][ncells] with the hydrogen moments of fields and k factors 1 (the
    context's source inputs are restored afterwards)."""
    f1 = np.array(fields, np.float64, copy=True)
    f1[16:19] = 1.0
    setka.set_sources(f1)
    q = analytic_sources(setka)
    setka.set_sources(fields)
    return q


def target_scale(setka, fields) -> np.ndarray:
    """Per-cell unit of the targets: src_coef * ro * n_H (n_H summed over populations)."""
    prim = setka.download()
    return setka.params.src_coef * prim[0] * fields[0:4].sum(axis=0)


OUT_SOURCE, OUT_FACTOR = 0, 1     # enum gel_mlp_output


class SourceModel:
    """The MLP plus its feature / target scaling.  mode OUT_SOURCE: outputs are the
    sources in units of src_coef * ro * n_H; OUT_FACTOR: factors on the closure without
    the Monte Carlo's k factors (source = closure x (1 + y))."""

    def __init__(self, mlp: MLP, mu, sd, ymu, ysd, mode: int = OUT_SOURCE):
        self.mlp = mlp
        self.mu = np.asarray(mu, np.float32)
        self.sd = np.asarray(sd, np.float32)
        self.ymu = np.asarray(ymu, np.float32)
        self.ysd = np.asarray(ysd, np.float32)
        self.mode = int(mode)
        L = _setup_src(_setup(_lib.lib()))
        _check(L.gel_mlp_set_scaling(mlp._h, _f(self.mu), _f(self.sd), _f(self.ymu), _f(self.ysd)))
        _check(L.gel_mlp_set_output_mode(mlp._h, self.mode))

    def predict(self, setka, fields) -> np.ndarray:
        """The network's sources [5][ncells] on a Setka's state (host-side de-normalisation,
        as the device refresh does it)."""
        y = self.mlp.forward((features(setka) - self.mu) / self.sd) * self.ysd + self.ymu
        if self.mode == OUT_FACTOR:
            return closure_k1(setka, fields) * (1.0 + y.T)
        return y.T * target_scale(setka, fields)

    def refresh(self, setka, tol: float = 1e-3, judge=True) -> int:
        """Fill the Setka's source table from the network (GEL_SRC_TABLE); returns the
        number of cells that fell back to the analytic source.  judge: False / 0 none,
        True / 1 every cell, 2 (GEL_SRC_HYBRID) only the cells not flagged already."""
        L = _setup_src(_setup(_lib.lib()))
        n = ctypes.c_longlong(0)
        _check(L.gel_surrogate_refresh(setka._ctx, self.mlp._h, float(tol), int(judge), ctypes.byref(n)))
        return n.value

    def advance(self, setka, steps: int, every: int = 10, tol: float = 1e-3):
        """Advance a GEL_SRC_HYBRID Setka `steps` steps with its sources refreshed from
        the network after every `every`-th step (sticky judge: cells already sent to the
        analytic closure keep it).  The first refresh, with every cell judged, is the
        caller's.  Returns the flagged cells after the last refresh (or None)."""
        done, nfall = 0, None
        while done < steps:
            n = min(every, steps - done)
            setka.Go_5_komponent__MK2(n)
            done += n
            if n == every:
                nfall = self.refresh(setka, tol, judge=2)
        return nfall

    def save(self, path: str) -> None:
        np.savez(path, params=self.mlp.get_params(), mu=self.mu, sd=self.sd, ymu=self.ymu, ysd=self.ysd,
                 mode=np.int32(self.mode))

    @classmethod
    def load(cls, path: str) -> "SourceModel":
        d = np.load(path, allow_pickle=False)
        m = MLP(0)
        m.set_params(d["params"])
        return cls(m, d["mu"], d["sd"], d["ymu"], d["ysd"], int(d["mode"]) if "mode" in d else OUT_SOURCE)


def inflow_sets(n_sets: int = 16, seed: int = 4):
    """Seeded inflow parameter sets: Mach uniform in [1.5, 3], n_H in [0.05, 0.2]."""
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(1.5, 3.0)), float(rng.uniform(0.05, 0.2)), seed * 1000 + k) for k in range(n_sets)]


def workload_with_hydrogen(ni: int, nj: int, steps: int, mach: float, nh: float, seed: int, moving: bool = True):
    """The 800 x 400 MHD workload's geometry and field with interstellar inflow Mach `mach` and hydrogen density
    `nh`; charge exchange with sigma(U) = (1 - 0.07 ln U)^2.  Returns (workload, fields)."""
    from .configs import R_BLEND, D_BLEND, mhd_heliosphere
    from .hydrogen import populations
    from .params import GEL_SRC_MK1
    w = mhd_heliosphere(ni, nj, steps, seed=seed, moving=moving)
    w.params.source_mode = GEL_SRC_MK1
    w.params.sig_a, w.params.sig_b = 1.0, 0.07
    u_inf = -w.params.inf_u
    p_inf = w.params.inf_rho * u_inf ** 2 / (w.params.gamma * mach ** 2)
    cx, cy = w.grid.cell_centers()
    s = 0.5 * (1.0 - np.tanh((np.hypot(cx, cy) - R_BLEND) / D_BLEND))
    w.prim0[1] = w.prim0[1] + (1 - s) * (p_inf - w.params.inf_p)
    w.params.inf_p = p_inf
    return w, populations(w.grid, n0=nh, u_inf=u_inf, seed=seed)


def training_data(ni: int = 500, nj: int = 250, n_sets: int = 16, seed: int = 4, max_evolve: int = 600):
    """Features and normalised-unit targets from the analytic closure on the device,
    over n_sets inflow parameter sets (ni * nj cells each); set k is sampled after
    k * max_evolve / n_sets steps of the sweep with analytic sources."""
    from .setka import Setka
    X, Y = [], []
    for k, (mach, nh, s) in enumerate(inflow_sets(n_sets, seed)):
        w, fields = workload_with_hydrogen(ni, nj, 0, mach, nh, s, moving=False)
        st = Setka(w.grid, w.params)
        st.upload(w.prim0)
        st.set_sources(fields)
        ev = k * max_evolve // max(n_sets, 1)
        if ev:
            st.Go_5_komponent__MK2(ev)
        x = features(st)
        q = analytic_sources(st)
        scale = target_scale(st, fields)
        X.append(x)
        Y.append((q / scale).T.astype(np.float32))
        st.close()
    return np.concatenate(X), np.concatenate(Y)


def mc_params_for(w, n_primary: int, seed: int):
    """Monte Carlo constants matching a workload's closure: the cross section
    (sig_a - sig_b ln U)^2 as (1 - a ln U)^2 with a = sig_b / sig_a, n_p / Kn = src_coef,
    the hydrogen drifting with the interstellar inflow."""
    from .params import mc_reference_params
    p = mc_reference_params(n_primary, seed=seed)
    p.sig_a = w.params.sig_b / w.params.sig_a
    p.kn = p.n_p_lism / w.params.src_coef
    p.v_inf = w.params.inf_u
    return p


def mc_sources(setka, w, nh: float, n_primary: int, seed: int):
    """The Monte Carlo (gel_mc_run) on a Setka's plasma, scaled to the hydrogen density
    nh at infinity: (19 source inputs, sources [5][ncells] = 0, I_u, I_v, I_T, 0, stats).
    The context (GEL_SRC_MK1) takes the 19 fields as its source inputs."""
    from .mc import run_mc
    out, st = run_mc(setka, mc_params_for(w, n_primary, seed))
    out[0:4] *= nh                  # densities and sources scale with the density at infinity
    out[19:25] *= nh
    setka.set_sources(out[:19])
    q = np.zeros((5, setka.ncells))
    q[1:4] = out[19:22]
    return out[:19].copy(), q, st


def training_data_mc(ni: int = 500, nj: int = 250, n_sets: int = 16, seed: int = 4, max_evolve: int = 600,
                     n_primary: int = 20_000_000, mode: int = OUT_FACTOR):
    """Features and targets from the hydrogen Monte Carlo on the device: per inflow set,
    the plasma after k * max_evolve / n_sets steps (analytic source from the seeded
    populations), the Monte Carlo's moments as the hydrogen features; targets its k
    factors less one (OUT_FACTOR: k = I / closure, Setka.cpp:12395-12449) or its momentum
    and energy sources I_u, I_v, I_T in units of src_coef * ro * n_H (OUT_SOURCE)."""
    from .setka import Setka
    X, Y = [], []
    for k, (mach, nh, s) in enumerate(inflow_sets(n_sets, seed)):
        w, fields = workload_with_hydrogen(ni, nj, 0, mach, nh, s, moving=False)
        st = Setka(w.grid, w.params)
        st.upload(w.prim0)
        st.set_sources(fields)
        ev = k * max_evolve // max(n_sets, 1)
        if ev:
            st.Go_5_komponent__MK2(ev)
        f, q, stats = mc_sources(st, w, nh, n_primary, s)
        print(f"inflow set {k}: Mach {mach:.2f} n_H {nh:.3f}, {ev} steps, Monte Carlo {stats['atoms']} atoms "
              f"in {stats['ms']:.0f} ms", flush=True)
        x = features(st)
        scale = target_scale(st, f)
        ok = scale > 0
        X.append(x[ok])
        if mode == OUT_FACTOR:
            y = np.zeros((5, st.ncells))
            y[1:4] = f[16:19] - 1.0
            Y.append(y[:, ok].T.astype(np.float32))
        else:
            Y.append((q[:, ok] / scale[ok]).T.astype(np.float32))
        st.close()
    return np.concatenate(X), np.concatenate(Y)


def train_source_model(X, Y, steps: int = 6000, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
                       shuffle_seed: int = 1, mode: int = OUT_SOURCE) -> SourceModel:
    """Normalise, shuffle, and train the MLP on the device; returns the model."""
    mu = X.mean(axis=0)
    sd = X.std(axis=0)
    sd = np.where(sd > 0, sd, 1.0)
    ymu = Y.mean(axis=0)
    ysd = Y.std(axis=0)
    const = ~(ysd > 0)
    ysd_train = np.where(const, 1.0, ysd)
    rng = np.random.default_rng(shuffle_seed)
    perm = rng.permutation(X.shape[0])
    Xn = ((X - mu) / sd)[perm].astype(np.float32)
    Yn = ((Y - ymu) / ysd_train)[perm].astype(np.float32)
    m = MLP(seed)
    m.set_dataset(Xn, Yn)
    done = 0
    history = []
    while done < steps:           # step the learning rate down in three stages
        k = min(steps - done, max(steps // 3, 1))
        stage = done * 3 // max(steps, 1)
        history.append(m.train(k, batch, lr * (0.3 ** stage)))
        done += k
    # constant targets are reproduced exactly by the device (ysd = 0)
    model = SourceModel(m, mu, sd, ymu, np.where(const, 0.0, ysd), mode)
    model.loss_history = np.concatenate(history)
    return model
