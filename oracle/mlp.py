"""The oracle (test infrastructure) for the sky-view surrogate MLP: an FP64 numpy
restatement of its forward pass and of the gradient of loss = scale * sum w (y - t)^2.

Two networks share the layers and the parameter layout:
- "f16_gate": hidden activation a(z) = z s(t), t = clamp(z / 6 + 1/2, 0, 1),
  s = t^2 (3 - 2 t), FP16 MFMA operands;
- "bf16_silu": hidden activation z sigmoid(z), BF16 MFMA operands.

With rounded=True every operand the GPU kernels feed to MFMA is rounded at the same points:
the inputs and the weights (to nearest even, FP16 or BF16), each hidden activation (for
f16_gate the kernels' packed FP16 sequence, emulated operation by operation: the FP32
pre-activation rounded toward zero, then fma, clamp, fma and three multiplies, each
rounded to nearest even; for bf16_silu the FP64 SiLU rounded to BF16), and each
back-propagated delta (kept without the loss scale, which multiplies the gradient sums).
Everything else, including the accumulations, runs in FP64.  Only tests/ and the
benchmark's checks use this module.
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


def f16_round(x):
    """Rounds to the nearest FP16 value (ties to even), returned as float64."""
    with np.errstate(over="ignore"):
        return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


def f16_round_toward_zero(x):
    """Rounds to FP16 toward zero (v_cvt_pkrtz_f16_f32), returned as float64; magnitudes
    beyond the largest finite value give that value."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(over="ignore"):
        r = np.atleast_1d(x.astype(np.float16))
    xa = np.atleast_1d(x)
    over = np.abs(r.astype(np.float64)) > np.abs(xa)
    r[over] = np.nextafter(r[over], np.float16(0.0))
    return r.astype(np.float64).reshape(x.shape)


SIXTH_F16 = float(np.float16(1.0 / 6.0))


def activation(z):
    t = np.clip(np.asarray(z, dtype=np.float64) / 6.0 + 0.5, 0.0, 1.0)
    return z * t * t * (3.0 - 2.0 * t)


def activation_grad(z):
    t = np.clip(np.asarray(z, dtype=np.float64) / 6.0 + 0.5, 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t) + z * t * (1.0 - t)


def activation_f16(z):
    """The kernels' packed FP16 evaluation of activation(z) for FP32 pre-activations z."""
    z = f16_round_toward_zero(np.asarray(z, dtype=np.float64).astype(np.float32))
    t = np.clip(f16_round(z * SIXTH_F16 + 0.5), 0.0, 1.0)
    v = f16_round(3.0 - 2.0 * t)
    w = f16_round(t * v)
    s = f16_round(t * w)
    return f16_round(z * s)


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


NETWORKS = ("f16_gate", "bf16_silu")


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def silu_grad(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def _network(network, rounded):
    """(operand rounding, activation, activation derivative) of a network."""
    if network not in NETWORKS:
        raise ValueError(network)
    exact = lambda v: np.asarray(v, dtype=np.float64)  # noqa: E731
    if network == "f16_gate":
        return (f16_round, activation_f16, activation_grad) if rounded else (exact, activation, activation_grad)
    return (bf16_round, lambda z: bf16_round(silu(z)), silu_grad) if rounded else (exact, silu, silu_grad)


def forward(params, x, network="f16_gate", rounded=False, cache=False):
    """x: [n, >= 7] inputs; returns y [n, 3] (and the pre-activations if cache)."""
    r, act, _ = _network(network, rounded)
    layers = unpack(params)
    a = r(np.asarray(x, dtype=np.float64)[:, :7]).T  # [7, n]
    zs, acts = [], [a]
    for li, (W, b) in enumerate(layers):
        z = r(W) @ a + b[:, None]
        if li == len(layers) - 1:
            return (z.T, zs, acts) if cache else z.T
        zs.append(z)
        a = act(z)
        acts.append(a)


def loss_and_grad(params, x, t, loss_scale, network="f16_gate", rounded=False, weights=None):
    """(loss, gradient as a flat vector) of loss = loss_scale * sum_n w_n sum_c (y - t)^2."""
    r, _, act_grad = _network(network, rounded)
    layers = unpack(params)
    y, zs, acts = forward(params, x, network=network, rounded=rounded, cache=True)
    d = y.T - np.asarray(t, dtype=np.float64).T  # [3, n]
    w = np.ones(d.shape[1]) if weights is None else np.asarray(weights, dtype=np.float64)
    loss = loss_scale * float(np.sum(w * d * d))
    delta = r(2.0 * w * d)  # without the loss scale
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        a_prev = acts[li]  # already rounded in forward when rounded
        grads[li] = (loss_scale * (delta @ a_prev.T), loss_scale * delta.sum(axis=1))
        if li > 0:
            delta = r((r(W).T @ delta) * act_grad(zs[li - 1]))
    return loss, pack(grads)
