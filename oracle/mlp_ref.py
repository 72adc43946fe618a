"""CPU reference of the surrogate network (float64 numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the GPU forward pass, the
GPU gradients (pelemp_mlp_loss_grad) and the bfloat16 MFMA inference.  The
reference has no learned component; this restates the network defined in
include/pelemp_spray.h (8 -> 4 x 64 hard swish -> 3, weights W_L[out][in] then
b_L per layer) and the loss
    L = (1/n) sum_i w_i |y_i - (t_i - out_mean) / out_std|^2
with plain reverse-mode differentiation.  With quant=True the weights and
each layer's input are rounded to bfloat16 (round to nearest even) in the
forward pass and the rounding is treated as the identity in the backward
pass (straight-through), as the GPU kernels do.

The backward pass is pinned by central finite differences in float64
(tests/test_mlp_oracle.py), the bfloat16 rounding by bit patterns.
"""
from __future__ import annotations

import numpy as np

FAN_IN = (8, 64, 64, 64, 64)
FAN_OUT = (64, 64, 64, 64, 3)
NPARAM = sum(i * o + o for i, o in zip(FAN_IN, FAN_OUT))


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray):
    p = np.asarray(params, dtype=np.float64)
    out, o = [], 0
    for i, n in zip(FAN_IN, FAN_OUT):
        W = p[o:o + i * n].reshape(n, i)
        o += i * n
        b = p[o:o + n]
        o += n
        out.append((W, b))
    assert o == NPARAM
    return out


def normalise_inputs(x: np.ndarray, center, scale) -> np.ndarray:
    # the GPU evaluates (x - c) * s in float32
    return ((np.asarray(x, np.float32) - np.asarray(center, np.float32)) * np.asarray(scale, np.float32)).astype(
        np.float64)


def act(z):
    """Hard swish z clamp(z/6 + 1/2, 0, 1), the hidden layers' activation."""
    return z * np.clip(z / 6.0 + 0.5, 0.0, 1.0)


def act_grad(z):
    return np.where(z <= -3.0, 0.0, np.where(z >= 3.0, 1.0, z / 3.0 + 0.5))


def forward(params, xn: np.ndarray, quant: bool = False):
    """Returns (y, cache) for normalised inputs xn (n x 8)."""
    layers = unpack(params)
    h = bf16_round(xn) if quant else np.asarray(xn, np.float64)
    cache = []
    for L, (W, b) in enumerate(layers):
        Wq = bf16_round(W) if quant else W
        z = h @ Wq.T + b
        cache.append((h, Wq, z))
        if L < len(layers) - 1:
            a = act(z)
            h = bf16_round(a) if quant else a
        else:
            h = z
    return h, cache


def loss_grad(params, xn, t, w, out_mean, out_std, quant: bool = False):
    """Loss and gradient (float64) of one batch."""
    n = xn.shape[0]
    y, cache = forward(params, xn, quant)
    tn = (np.asarray(t, np.float64) - np.asarray(out_mean, np.float64)) / np.asarray(out_std, np.float64)
    w = np.ones(n) if w is None else np.asarray(w, np.float64)
    e = y - tn
    loss = float(np.sum(w[:, None] * e * e) / n)
    delta = 2.0 * w[:, None] * e / n
    grads = []
    for L in range(len(cache) - 1, -1, -1):
        h, Wq, z = cache[L]
        if L < len(cache) - 1:
            delta = delta * act_grad(z)
        grads.append((delta.T @ h, delta.sum(axis=0)))
        delta = delta @ Wq
    grads.reverse()
    g = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, g
