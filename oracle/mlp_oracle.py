"""NumPy restatement of the surrogate network (14 -> 4 x 128 SiLU -> 5).

TEST INFRASTRUCTURE: the checker for the MFMA forward and backward kernels of
gelios_2d_amd/csrc/surrogate.hip.  Two modes:
  - exact: FP64 everywhere (pinned by finite differences in tests);
  - bf16:  the device's rounding points -- inputs, weights, activations and the
    back-propagated pre-activation gradients rounded to BF16 (round to nearest even),
    products accumulated in FP32 -- so that device results can be compared closely.
Parameter packing: for each layer, W [out][in] row-major then b [out].
"""
from __future__ import annotations

import numpy as np

SIZES = [(14, 128), (128, 128), (128, 128), (128, 128), (128, 5)]


def n_params() -> int:
    return sum(i * o + o for i, o in SIZES)


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    out, o = [], 0
    for i, n in SIZES:
        W = np.asarray(p[o:o + i * n]).reshape(n, i)
        o += i * n
        b = np.asarray(p[o:o + n])
        o += n
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, mode="exact"):
    L = unpack(p)
    if mode == "exact":
        h = np.asarray(X, np.float64)
        zs, hs = [], [h]
        for k, (W, b) in enumerate(L):
            z = h @ np.asarray(W, np.float64).T + b
            zs.append(z)
            h = _silu(z) if k < 4 else z
            hs.append(h)
        return h, (zs, hs)
    h = bf16(X).astype(np.float32)
    zs, hs = [], [h]
    for k, (W, b) in enumerate(L):
        z = (h.astype(np.float32) @ bf16(W).T.astype(np.float32)).astype(np.float32) + np.float32(1.0) * b.astype(np.float32)
        zs.append(z)
        h = bf16(_silu(z).astype(np.float32)) if k < 4 else z
        hs.append(h)
    return h, (zs, hs)


def loss_and_grad(p, X, T, mode="exact"):
    """Mean squared error over the n x 5 outputs and its gradient (packed)."""
    L = unpack(p)
    Y, (zs, hs) = forward(p, X, mode)
    n = X.shape[0]
    e = Y - T
    loss = float(np.mean(e.astype(np.float64) ** 2))
    d = 2.0 * e / (n * 5)
    if mode != "exact":
        d = bf16(d.astype(np.float32))
    grads = [None] * 5
    for k in range(4, -1, -1):
        W, b = L[k]
        hprev = hs[k]
        gW = d.T @ hprev
        gb = d.sum(axis=0)
        grads[k] = (gW, gb)
        if k > 0:
            dh = d @ (W if mode == "exact" else bf16(W))
            d = dh * _dsilu(zs[k - 1])
            if mode != "exact":
                d = bf16(d.astype(np.float32))
    return loss, pack(grads)
