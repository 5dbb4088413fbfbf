"""CPU restatement of the learned warm-start network (ORACLE: test infrastructure).

Only tests/ import this module.  It recomputes, in numpy, what the GPU's bf16
MFMA kernels compute: a 6 -> 32 -> 32 -> 32 -> 2 MLP with SiLU, weights and
layer inputs rounded to bfloat16 (round to nearest even) where the GPU feeds
them to the matrix cores, products accumulated in float32 (the GPU sums in a
different order), biases and pre-activations in float32.

Flat parameter vector (2402 floats): W1[32][6] b1[32] W2[32][32] b2[32]
W3[32][32] b3[32] W4[2][32] b4[2].  norm = (mu_in[6], sd_in[6], mu_out[2], sd_out[2]).

`loss_grad` differentiates the mean over particles of the squared error of
the normalised outputs, summed over the two outputs, as the training kernel
does; with `exact=True` no rounding is applied (float64), which is what the
finite-difference test of this module checks.
"""
from __future__ import annotations

import numpy as np

SIZES = [(32, 6), (32, 32), (32, 32), (2, 32)]
NP = sum(o * i + o for o, i in SIZES)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32)
    r = ((u >> 16) & 1) + np.uint32(0x7FFF)
    return ((u + r) & np.uint32(0xFFFF0000)).view(np.float32)


def unpack(P):
    out, o = [], 0
    for m, n in SIZES:
        W = np.asarray(P[o:o + m * n]).reshape(m, n)
        o += m * n
        b = np.asarray(P[o:o + m])
        o += m
        out.append((W, b))
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_d(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(P, norm, feat, exact=False):
    """Returns (y_normalised [n,2], cache)."""
    dt = np.float64 if exact else np.float32
    rnd = (lambda a: np.asarray(a, dt)) if exact else bf16
    norm = np.asarray(norm, dt)
    mu_in, sd_in = norm[0:6], norm[6:12]
    x = (np.asarray(feat, dt) - mu_in) * (dt(1.0) / sd_in)
    h = rnd(x)
    layers = unpack(np.asarray(P, dt))
    hs, zs = [h], []
    for L, (W, b) in enumerate(layers):
        z = (h.astype(dt) @ rnd(W).astype(dt).T) + b
        zs.append(z)
        if L < 3:
            h = rnd(_silu(z).astype(dt))
            hs.append(h)
    return zs[-1], dict(hs=hs, zs=zs, layers=layers, rnd=rnd, dt=dt)


def predict(P, norm, feat, exact=False):
    y, _ = forward(P, norm, feat, exact)
    norm = np.asarray(norm, np.float64)
    return y * norm[14:16] + norm[12:14]


def _backward(cache, g):
    """g: dL/dy (normalised units) [n,2]; returns (dL/dP flat, dL/d(normalised input) [n,6])."""
    hs, zs, layers, rnd, dt = cache["hs"], cache["zs"], cache["layers"], cache["rnd"], cache["dt"]
    grads = [None] * 4
    dz = g
    dx = None
    for L in range(3, -1, -1):
        W, _ = layers[L]
        dzr = rnd(dz)
        dW = dzr.astype(dt).T @ hs[L].astype(dt)
        db = dz.sum(axis=0)
        grads[L] = (dW, db)
        dh = dzr.astype(dt) @ rnd(W).astype(dt)
        if L > 0:
            dz = dh * _silu_d(zs[L - 1])
        else:
            dx = dh
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return flat, dx


def loss_grad(P, norm, feat, targ, exact=False):
    """Sum over particles of the squared normalised error (as the kernel's loss
    accumulator), and the gradient of its mean (loss / n) with respect to P."""
    y, cache = forward(P, norm, feat, exact)
    dt = cache["dt"]
    norm = np.asarray(norm, dt)
    t = (np.asarray(targ, dt) - norm[12:14]) / norm[14:16]
    e = y - t
    n = y.shape[0]
    g = (dt(2.0) * e) * dt(1.0 / n)
    flat, _ = _backward(cache, g)
    return float(np.sum(np.asarray(e, np.float64) ** 2)), flat


def input_grad(P, norm, feat, ybar, exact=False):
    """dL/d(raw features) for dL/d(raw predictions) = ybar."""
    y, cache = forward(P, norm, feat, exact)
    dt = cache["dt"]
    norm = np.asarray(norm, dt)
    g = np.asarray(ybar, dt) * norm[14:16]
    _, dx = _backward(cache, g)
    return dx / norm[6:12]
