"""CPU restatement of the radiance-cache MLP (numpy).

TEST INFRASTRUCTURE: the checker for the GPU cache kernels.  Same encoding,
layer shapes, padding, loss and parameter layout as the GPU; optional
emulation of the GPU's BF16 roundings (inputs, weights, stored activations,
back-propagated gradients), with sums in float64.

Pinned by: its backward pass against central finite differences of its own
float64 forward pass (tests/test_nrc.py), and the encoding against its
closed form (sin / cos of 2^k pi x).
"""
from __future__ import annotations

import numpy as np

KD = [96, 64, 64, 64, 64, 64]
ND = [64, 64, 64, 64, 64, 16]
K0_VALID, N5_VALID, FREQS = 78, 3, 12


def _offsets():
    ow, ob, o = [], [], 0
    for k, n in zip(KD, ND):
        ow.append(o)
        o += n * k
        ob.append(o)
        o += n
    return ow, ob, o


OFFW, OFFB, NPARAM = _offsets()


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


def unpack(params: np.ndarray):
    W = [params[OFFW[l]:OFFW[l] + ND[l] * KD[l]].reshape(ND[l], KD[l]) for l in range(6)]
    b = [params[OFFB[l]:OFFB[l] + ND[l]] for l in range(6)]
    return W, b


def encode(feats: np.ndarray, lo, inv) -> np.ndarray:
    """(n, 9) -> (n, 96): per axis a and k < 12, sin / cos(2^k pi x_a) with x
    the position normalised to the box; then normal, direction, zeros."""
    feats = np.asarray(feats, np.float32)
    n = feats.shape[0]
    X = np.zeros((n, 96), np.float64)
    lo = np.asarray(lo, np.float32)
    inv = np.asarray(inv, np.float32)
    for a in range(3):
        x = ((feats[:, a] - lo[a]) * inv[a]).astype(np.float32)
        for k in range(FREQS):
            arg = np.ldexp(x, k).astype(np.float64)
            X[:, a * 24 + 2 * k] = np.sin(np.pi * np.mod(arg, 2.0))
            X[:, a * 24 + 2 * k + 1] = np.cos(np.pi * np.mod(arg, 2.0))
    X[:, 72:78] = feats[:, 3:9]
    return X


def forward(params, feats, lo, inv, emulate: bool = True):
    """Returns (y (n, 3), activations [X, H0..H4], pre-activation outputs)."""
    W, b = unpack(np.asarray(params, np.float32))
    r = bf16 if emulate else (lambda a: a)
    X = encode(feats, lo, inv)
    X = r(X.astype(np.float32)).astype(np.float64) if emulate else X
    acts = [X]
    h = X
    for l in range(5):
        z = h @ r(W[l]).astype(np.float64).T + b[l]
        h = np.maximum(z, 0.0)
        if emulate:
            h = r(h.astype(np.float32)).astype(np.float64)
        acts.append(h)
    y = h @ r(W[5]).astype(np.float64).T + b[5]
    return y[:, :N5_VALID], acts


def loss_and_grad(params, feats, targets, lo, inv, emulate: bool = True):
    """Mean relative L2 loss (y - t)^2 / (y^2 + 0.01) over n x 3 outputs, and
    the gradient of sum over samples of that per-sample loss / n (the GPU's
    scaling), in the parameter layout."""
    params = np.asarray(params, np.float32)
    W, b = unpack(params)
    r = bf16 if emulate else (lambda a: a)
    y, acts = forward(params, feats, lo, inv, emulate)
    t = np.asarray(targets, np.float64)
    n = y.shape[0]
    den = y * y + 0.01
    loss = float(np.sum((y - t) ** 2 / den) / (n * N5_VALID))
    dy = np.zeros((n, 16))
    dy[:, :N5_VALID] = 2.0 * (y - t) / den / n
    if emulate:
        dy = r(dy.astype(np.float32)).astype(np.float64)
    g = np.zeros(NPARAM, np.float64)
    dz = dy
    for l in range(5, -1, -1):
        inp = acts[l]
        g[OFFW[l]:OFFW[l] + ND[l] * KD[l]] = (dz.T @ inp).reshape(-1)
        g[OFFB[l]:OFFB[l] + ND[l]] = dz.sum(axis=0)
        if l > 0:
            dh = dz @ r(W[l]).astype(np.float64)
            dz = np.where(acts[l] > 0, dh, 0.0)
            if emulate:
                dz = r(dz.astype(np.float32)).astype(np.float64)
    return loss, g


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights on the valid inputs / outputs, zero biases and
    padding (same scheme as the GPU's initialisation, different generator)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for l in range(6):
        kv = K0_VALID if l == 0 else KD[l]
        nv = N5_VALID if l == 5 else ND[l]
        a = np.sqrt(6.0 / kv)
        w = np.zeros((ND[l], KD[l]), np.float32)
        w[:nv, :kv] = rng.uniform(-a, a, (nv, kv))
        p[OFFW[l]:OFFW[l] + ND[l] * KD[l]] = w.reshape(-1)
    return p
