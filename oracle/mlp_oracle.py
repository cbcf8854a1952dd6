"""FP64 restatement of the per-particle MLP (8 -> 64 -> 64 -> 64 -> 1, tanh-form
GELU) used by the learned warm start, with its backward pass.

TEST INFRASTRUCTURE ONLY (the checker of the HIP forward/backward kernels).
`bf16=True` rounds to bfloat16 at the points where the GPU feeds its matrix
cores (inputs, weights, hidden activations, back-propagated gradients); all
arithmetic is FP64.  The network has no counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

SIZES = [(64, 8), (64, 64), (64, 64), (1, 64)]   # (out, in) per layer
N_PARAMS = sum(o * i + o for o, i in SIZES)     # 8961
C_GELU = np.sqrt(2.0 / np.pi)


def bf16(a):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    x = np.asarray(a, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, k = [], 0
    for o, i in SIZES:
        W = p[k:k + o * i].reshape(o, i)
        k += o * i
        b = p[k:k + o]
        k += o
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * x * x)


def forward(p, feat, bf=False):
    """Returns y (n,) and the cache of pre-activations."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers = unpack(p)
    h = r(feat).T                      # in x n
    zs = []
    for L, (W, b) in enumerate(layers):
        z = r(W) @ h + b[:, None]
        if L < 3:
            zs.append(z)
            h = r(gelu(z))
        else:
            y = z[0]
    return y, (zs, r(feat).T)


def gradient(p, feat, target, bf=False, over=1.0):
    """Gradient of mean(w (y - t)^2), w = `over` where y > t else 1, and the loss."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers = unpack(p)
    n = feat.shape[0]
    y, (zs, x) = forward(p, feat, bf)
    e = y - np.asarray(target, np.float64)
    w = np.where(e > 0, over, 1.0)
    loss = float(np.mean(w * e * e))
    dz = (2.0 * w * e / n)[None, :]    # 1 x n
    hs = [x] + [r(gelu(z)) for z in zs]
    grads = [None] * 4
    for L in range(3, -1, -1):
        W, _ = layers[L]
        dW = r(dz) @ hs[L].T
        db = dz.sum(axis=1)
        grads[L] = (dW, db)
        if L > 0:
            dh = r(W).T @ r(dz)
            dz = dh * gelu_grad(zs[L - 1])
    return pack(grads), loss


def loss_only(p, feat, target):
    y, _ = forward(p, feat, False)
    e = y - np.asarray(target, np.float64)
    return float(np.mean(e * e))


def glorot(seed=0):
    rng = np.random.default_rng(seed)
    layers = []
    for o, i in SIZES:
        a = np.sqrt(6.0 / (i + o))
        layers.append((rng.uniform(-a, a, (o, i)), np.zeros(o)))
    return pack(layers)
