"""The oracle (test infrastructure) for the sky-view surrogate MLP: an FP64 numpy
restatement of its forward pass and of the gradient of loss = scale * sum (y - t)^2.

With bf16=True every operand the GPU kernels feed to BF16 MFMA is rounded to BF16
(round to nearest even) at the same points: the inputs, the weights, each hidden
activation, and each backpropagated delta (the deltas before they enter the weight
gradient, the bias sums and the next delta product).  Everything else, including the
accumulations, runs in FP64.  Only tests/ and the benchmark's checks use this module.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(7, 64), (64, 64), (64, 64), (64, 64), (64, 3)]  # (fan-in, fan-out)
PARAMS = sum(k * n + n for k, n in LAYERS)


def bf16_round(x):
    """Rounds to the nearest BF16 value (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    """[(W (fan_out x fan_in), b (fan_out))] from the flat parameter vector."""
    p = np.asarray(params, dtype=np.float64)
    out, o = [], 0
    for k, n in LAYERS:
        W = p[o:o + k * n].reshape(n, k)
        o += k * n
        b = p[o:o + n]
        o += n
        out.append((W, b))
    assert o == PARAMS
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def silu_grad(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, bf16=False, cache=False):
    """x: [n, >= 7] inputs; returns y [n, 3] (and the pre-activations if cache)."""
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params)
    a = r(np.asarray(x, dtype=np.float64)[:, :7]).T  # [7, n]
    zs, acts = [], [a]
    for li, (W, b) in enumerate(layers):
        z = r(W) @ a + b[:, None]
        if li == len(layers) - 1:
            return (z.T, zs, acts) if cache else z.T
        zs.append(z)
        a = r(silu(z))
        acts.append(a)


def loss_and_grad(params, x, t, loss_scale, bf16=False, weights=None):
    """(loss, gradient as a flat vector) of loss = loss_scale * sum_n w_n sum_c (y - t)^2."""
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params)
    y, zs, acts = forward(params, x, bf16=bf16, cache=True)
    d = y.T - np.asarray(t, dtype=np.float64).T  # [3, n]
    w = np.ones(d.shape[1]) if weights is None else np.asarray(weights, dtype=np.float64)
    loss = loss_scale * float(np.sum(w * d * d))
    delta = r(2.0 * loss_scale * w * d)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        a_prev = acts[li]  # already rounded in forward when bf16
        grads[li] = (delta @ a_prev.T, delta.sum(axis=1))
        if li > 0:
            delta = r((r(W).T @ delta) * silu_grad(zs[li - 1]))
    return loss, pack(grads)
