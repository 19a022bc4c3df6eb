"""Pins the surrogate's oracle (oracle/mlp.py) on the CPU.

- FP16 and BF16 rounding (to nearest even) against torch's float32 conversions; rounding
  toward zero against its definition (the FP16 value of largest magnitude not above |x|,
  found by stepping through the FP16 grid) and the range limit.
- The activation: its closed form at points where t is 0, 1/2 and 1, its derivative against
  central differences, and the FP16 operation sequence against the FP64 function within
  a few FP16 ulps.
- The analytic gradient against central finite differences in FP64 (no operand rounding,
  which would make the loss piecewise constant), for every layer's weights and biases, for
  both networks.
- A one-hidden-unit special case in closed form: y = w2 a(a(a(a(w1 x + b1)))) + b2, for the
  gate and for SiLU (against math.exp).
- The FP64 sky-radiance query against the sky-view texel path it restates: a query
  along a sky-view texel's direction equals the oracle's sky-view LUT texel.
"""
import math

import numpy as np
import pytest
import torch

import oracle
from oracle import mlp
from skyrendering_amd import presets
from skyrendering_amd.atmosphere import AtmosphereRenderParameters


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.uniform(-8, 8, 20000),
                        np.array([0.0, 1.0, -1.0, 1.00390625, 1.005859375, 3.0e38])]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = mlp.bf16_round(x).astype(np.float32)
    assert np.array_equal(got, want)


def test_f16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.uniform(-8, 5, 20000),
                        np.array([0.0, 1.0, -1.0, 1.00048828125, 1.00146484375, 65504.0, 6.0e-8])]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.float16).to(torch.float32).numpy()
    got = mlp.f16_round(x).astype(np.float32)
    assert np.array_equal(got, want)


def test_f16_round_toward_zero():
    rng = np.random.default_rng(2)
    x = rng.standard_normal(5000) * 10.0 ** rng.uniform(-6, 4, 5000)
    got = mlp.f16_round_toward_zero(x)
    grid = np.sort(np.unique(np.arange(0, 0x7C00, dtype=np.uint16).view(np.float16).astype(np.float64)))
    want = np.sign(x) * grid[np.searchsorted(grid, np.abs(x), side="right") - 1]
    assert np.array_equal(got, want)
    assert np.array_equal(mlp.f16_round_toward_zero(np.array([7e4, -1e9, 1.0, -0.0])), [65504.0, -65504.0, 1.0, 0.0])


def test_activation_closed_form_and_derivative():
    a, g = mlp.activation, mlp.activation_grad
    assert a(-3.0) == 0.0 and a(-10.0) == 0.0 and a(0.0) == 0.0 and a(3.0) == 3.0 and a(7.5) == 7.5
    assert a(1.5) == 1.5 * (0.75 ** 2 * 1.5)  # t = 3/4
    z = np.linspace(-4.0, 4.0, 81) + 0.013
    h = 1e-6
    assert np.allclose(g(z), (a(z + h) - a(z - h)) / (2 * h), rtol=0, atol=1e-8)
    z = np.random.default_rng(3).uniform(-8, 8, 20000)
    err = np.abs(mlp.activation_f16(z) - a(z))
    assert np.all(err <= 4 * 2.0 ** -10 * np.maximum(np.abs(z), 2.0 ** -4))


@pytest.mark.parametrize("network", mlp.NETWORKS)
def test_gradient_matches_finite_differences(network):
    rng = np.random.default_rng(1)
    p = rng.standard_normal(mlp.PARAMS) * 0.3
    x = rng.uniform(-1, 1, size=(9, 8))
    t = rng.standard_normal((9, 3))
    loss, g = mlp.loss_and_grad(p, x, t, 0.1, network=network)
    assert abs(loss - 0.1 * np.sum((mlp.forward(p, x, network=network) - t) ** 2)) < 1e-12
    idx = np.concatenate([rng.choice(np.arange(o, o + k * n + n), 12, replace=False)
                          for o, (k, n) in zip(np.cumsum([0] + [k * n + n for k, n in mlp.LAYERS[:-1]]), mlp.LAYERS)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss_and_grad(p + e, x, t, 0.1, network=network)[0] -
              mlp.loss_and_grad(p - e, x, t, 0.1, network=network)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


@pytest.mark.parametrize("network", mlp.NETWORKS)
def test_single_path_closed_form(network):
    """Only unit 0 of each hidden layer active with unit weights: the network reduces to
    nested activations of the first input, written out here."""
    layers = [(np.zeros((n, k)), np.zeros(n)) for k, n in mlp.LAYERS]
    layers[0][0][0, 0] = 1.5
    layers[0][1][0] = 0.25
    for li in (1, 2, 3):
        layers[li][0][0, 0] = 1.0
    layers[4][0][:, 0] = [1.0, 2.0, -1.0]
    layers[4][1][:] = [0.5, 0.0, 0.0]
    p = mlp.pack(layers)
    def s(z):
        if network == "bf16_silu":
            return z / (1.0 + math.exp(-z))
        t = min(max(z / 6.0 + 0.5, 0.0), 1.0)
        return z * t * t * (3.0 - 2.0 * t)
    for x0 in (-2.0, 0.3, 1.7):
        h = s(s(s(s(1.5 * x0 + 0.25))))
        y = mlp.forward(p, np.array([[x0, 0, 0, 0, 0, 0, 0, 0]]), network=network)[0]
        assert np.allclose(y, [h + 0.5, 2 * h, -h], rtol=1e-14, atol=1e-15)


def test_sky_radiance_query_equals_sky_view_texel():
    atm_p = presets.earth_100km()
    atm = atm_p.buffer_data()
    T = oracle.transmittance_lut(atm, 256, 64).astype(np.float32)
    MS = oracle.multiscattering_lut(atm, T, 32, 32).astype(np.float32)
    cam = presets.bench_camera()
    theta = 63.0
    rp = AtmosphereRenderParameters(sun_direction_theta=theta, sun_direction_phi=180.0, sky_view_lut_steps=30.0,
                                    camera_position=cam.position, view_projection=cam.view_projection())
    render = rp.buffer_data(atm_p.bottom_radius)
    L, _ = oracle.sky_view_lut(atm, render, T, MS, 24, 16)
    # Texel (ix, iy): the oracle's latitude / longitude parameterisation, restated here.
    r = float(render.camera_earth_center_distance)
    rb = float(atm.bottom_radius)
    down = math.acos(math.sqrt(r * r - rb * rb) / r)
    up = math.pi - down
    mu_s = float(np.dot(np.array(render.sun_direction[:], np.float64), np.array(render.up_direction[:], np.float64)))
    queries, texels = [], []
    for ix, iy in [(3, 2), (11, 5), (20, 7), (0, 9), (17, 12), (23, 15)]:
        xl, yl = ix / 23, iy / 15
        lat = up * (1 - (1 - 2 * yl) ** 2) if yl < 0.5 else up + down * (2 * yl - 1) ** 2
        cos_lon = -(xl * xl * 2 - 1)
        queries.append([0, r - rb, math.cos(lat), mu_s, math.acos(max(-1.0, min(1.0, cos_lon)))])
        texels.append(L[iy, ix, :3])
    q = np.ascontiguousarray(np.array(queries, np.float64))
    out = np.empty((len(q), 3), np.float64)
    atms = (type(atm) * 1)(atm)
    oracle.lib().oracle_sky_radiance_f64(atms, T.ctypes.data, 256, 64, MS.ctypes.data, 32, 32, 30.0, q.ctypes.data,
                                         len(q), out.ctypes.data)
    assert np.allclose(out, np.array(texels), rtol=2e-6, atol=0), (out, texels)
