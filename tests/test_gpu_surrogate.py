"""GPU checks of the learned sky-view surrogate against the oracle.

- Inference (register-resident weights; both networks, FP16 MFMA with packed FP16 gate
  activations and BF16 MFMA with SiLU): against the FP64 oracle forward with the same
  operand rounding and activation; an FP32 pre-activation that lands next to a rounding
  boundary can round differently than the oracle's FP64 sum (and SiLU's hardware exp and
  reciprocal differ from FP64 by an ulp), which moves one activation by one ulp, so at
  least 99 % of outputs agree within 1e-3 and all within 5e-2.
- Gradient (training kernel, MFMA forward + backward, fixed-order reduction; both
  networks): against the FP64 oracle gradient with the same rounding, per layer relative
  L2 < 1e-2 and cosine similarity > 0.9999; against the unrounded FP64 gradient the
  rounding effect is reported.
- Labels (plain integrator per query): against the FP64 oracle, 1e-3 relative.
- Training reduces the loss; the render pipeline's fallback recomputes exactly the flagged
  tiles with the plain integrator, and a trained network serves most tiles within the bound.
- The render pass (64 pixels per wave staged through LDS) gives every pixel the network
  value that the feature and forward entry points give for that pixel's query.
- The feature entry point against its definition (the inputs have no reference
  counterpart: their formulas, restated here in FP64, are the specification).
"""
import math

import numpy as np
import pytest

import oracle
from oracle import mlp
from skyrendering_amd import surrogate as sg

pytestmark = pytest.mark.gpu


def _params(seed=3, scale=2.0):
    p = sg.init_params(seed).astype(np.float64)
    p *= scale  # larger weights: pre-activations spread over the activation's curved range
    p[13184:13187] = [-3.0, -2.0, -1.0]
    return p.astype(np.float32)


def _features(n, seed=4):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 8), np.float32)
    f[:, :7] = rng.uniform(-1, 1, size=(n, 7))
    return f


@pytest.mark.parametrize("network", ["f16_gate", "bf16_silu"])
def test_forward_matches_oracle(ctx, network):
    p = _params()
    f = _features(1000)
    model = sg.SkySurrogate(ctx, p, network=network)
    out = ctx.alloc(1000 * 12)
    model.forward(ctx.upload(f), 1000, out)
    got = out.download(np.float32, (1000, 3)).astype(np.float64)
    ref = mlp.forward(p.astype(np.float64), f.astype(np.float64), network=network, rounded=True)
    err = np.abs(got - ref) / (1.0 + np.abs(ref))
    ref_fp64 = mlp.forward(p.astype(np.float64), f.astype(np.float64), network=network)
    e64 = np.abs(got - ref_fp64) / (1.0 + np.abs(ref_fp64))
    print(f"{network}: forward vs the rounded oracle: median {np.median(err):.2e}, p99 {np.quantile(err, 0.99):.2e}, "
          f"max {err.max():.2e}; vs unrounded FP64: median {np.median(e64):.2e}, max {e64.max():.2e}")
    assert np.mean(err <= 1e-3) >= 0.99 and err.max() < 5e-2


@pytest.mark.parametrize("network", ["f16_gate", "bf16_silu"])
def test_gradient_matches_oracle(ctx, network):
    p = _params(seed=5, scale=1.0)
    n = 96  # three training workgroups
    f = _features(n, seed=6)
    t = np.random.default_rng(7).normal(-2.0, 1.0, size=(n, 3)).astype(np.float32)
    w = np.random.default_rng(8).uniform(0.1, 1.0, n).astype(np.float32)
    model = sg.SkySurrogate(ctx, p, network=network)
    fd, td, wd = ctx.upload(f), ctx.upload(t), ctx.upload(w)
    scale = 1.0 / (3 * n)
    model.gradient(fd.ptr, td.ptr, n, scale, wd.ptr)
    g = model.grad.download(np.float32, (sg.PARAMS,)).astype(np.float64)
    loss = float(model.loss.download(np.float32, (1,))[0])
    ref_loss, ref = mlp.loss_and_grad(p.astype(np.float64), f.astype(np.float64), t.astype(np.float64), scale,
                                      network=network, rounded=True, weights=w)
    _, ref64 = mlp.loss_and_grad(p.astype(np.float64), f.astype(np.float64), t.astype(np.float64), scale,
                                 network=network, weights=w)
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    o = 0
    for li, (k, nn) in enumerate(mlp.LAYERS):
        for name, size in (("W", k * nn), ("b", nn)):
            a, b, c = g[o:o + size], ref[o:o + size], ref64[o:o + size]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
            rel64 = np.linalg.norm(a - c) / np.linalg.norm(c)
            print(f"{network} layer {li + 1} {name}: rel L2 {rel:.2e} (vs unrounded FP64 {rel64:.2e}), cosine {cos:.8f}")
            assert rel < 1e-2 and cos > 0.9999
            o += size


def _small_atmospheres(ctx, n=4, seed=9):
    return sg.SkyAtmospheres(ctx, sg.draw_scales(n, seed))


def test_labels_match_oracle(ctx):
    atm = _small_atmospheres(ctx)
    q = sg.random_queries(3000, 4, atm.thickness, seed=10)
    qd = ctx.upload(q)
    out = ctx.alloc(3000 * 12)
    atm.radiance(qd, 3000, out)
    got = out.download(np.float32, (3000, 3)).astype(np.float64)
    blocks, T, MS = atm.host_luts()
    qq = np.stack([q["atmosphere"], q["altitude"], q["mu_view"], q["mu_sun"], q["delta_phi"]], 1).astype(np.float64)
    ref = oracle.sky_radiance(blocks, T, MS, 30.0, qq)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6 * np.abs(ref).max())
    print(f"labels: max rel {err.max():.2e}, median {np.median(err):.2e}")
    assert err.max() < 1e-3


@pytest.mark.parametrize("network", ["f16_gate", "bf16_silu"])
def test_training_reduces_loss(ctx, network):
    atm = sg.SkyAtmospheres(ctx, sg.draw_scales(16, 11))
    data = sg.SkyTrainingSet(ctx, atm, 1 << 16, seed=12)
    model = sg.SkySurrogate(ctx, seed=13, network=network)
    model.set_output_bias(data.targets_host.mean(axis=0))
    hist = model.train(data, epochs=6, batch=4096, lr=3e-3, lr_final=3e-4)
    print("loss per epoch:", [f"{h:.3e}" for h in hist])
    assert hist[-1] < 0.2 * hist[0]
    # The inference kernel agrees with the training kernel's loss on the training data.
    out = ctx.alloc(data.n * 12)
    model.forward(data.features, data.n, out)
    y = out.download(np.float32, (data.n, 3))
    mse = float(np.mean(data.weights_host[:, None] * (y - data.targets_host) ** 2))
    print(f"inference weighted MSE {mse:.3e}, last epoch loss {hist[-1]:.3e}")
    assert mse < 2.0 * hist[-1]


def test_render_pipeline_fallback(ctx):
    scales, views = sg.evaluation_views(3, seed=14)
    atm = sg.SkyAtmospheres(ctx, scales, views=views)
    pano = sg.SkyPanoramas(ctx, atm, views, 160, 72)
    pano.render_plain()
    plain = pano.image()
    model = sg.SkySurrogate(ctx, seed=15)  # untrained: every tile must fall back
    pano.render_surrogate(model, tolerance=1e-2)
    assert pano.fallback_tiles() == pano.tiles * 3
    assert np.array_equal(pano.image(), plain)
    pano.render_surrogate(model, tolerance=1e30)  # nothing falls back: network values except probes
    assert pano.fallback_tiles() == 0
    img = pano.image()
    assert not np.array_equal(img, plain) and np.all(np.isfinite(img))
    # Sky-view LUT sampling, the method without the surrogate, approximates the plain pass.
    pano.render_lut()
    lut = pano.image()
    e = sg.relative_error(lut[..., :3].astype(np.float64), plain[..., :3].astype(np.float64), 1e-6)
    print(f"sky-view LUT vs plain: median {np.median(e):.2e}, p99 {np.quantile(e, 0.99):.2e}")
    assert np.median(e) < 2e-2
    # Plain pass against the FP64 oracle on sampled pixels.
    blocks, T, MS = atm.host_luts()
    rng = np.random.default_rng(16)
    xs, ys = rng.integers(0, 160, 200), rng.integers(0, 72, 200)
    ref = oracle.sky_radiance(blocks, T, MS, 30.0, pano.pixel_queries(1, xs, ys))
    e = sg.relative_error(plain[1, ys, xs, :3].astype(np.float64), ref, 1e-6)
    print(f"plain pass vs FP64 oracle: max {e.max():.2e}")
    assert e.max() < 1e-3


@pytest.mark.parametrize("network", ["f16_gate", "bf16_silu"])
def test_render_pass_matches_forward_entry_points(ctx, network):
    scales, views = sg.evaluation_views(2, seed=16)
    atm = sg.SkyAtmospheres(ctx, scales, views=views)
    w, h = 200, 90  # 18,000 pixels per view: the last 64-pixel run of the last view is partial
    pano = sg.SkyPanoramas(ctx, atm, views, w, h)
    model = sg.SkySurrogate(ctx, _params(seed=17, scale=1.0), network=network)
    pano.render_surrogate(model, tolerance=1e30)
    img = pano.image()[..., :3].astype(np.float64)
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    probe = np.zeros((h, w), bool)  # the residual check leaves the plain value at its probes
    for ty in range(0, h, 16):
        for tx in range(0, w, 16):
            for i in range(4):
                probe[min(ty + 4 + 8 * (i >> 1), h - 1), min(tx + 4 + 8 * (i & 1), w - 1)] = True
    for v in range(2):
        q = pano.pixel_queries(v, xs.ravel(), ys.ravel())
        qs = np.zeros(len(q), sg.QUERY_DTYPE)
        for k, name in enumerate(("atmosphere", "altitude", "mu_view", "mu_sun", "delta_phi")):
            qs[name] = q[:, k]
        n = len(qs)
        feats = ctx.alloc(n * 32)
        from skyrendering_amd import _abi
        _abi.check(_abi.lib().skyhip_surrogate_features(ctx.handle, ctx.upload(qs).ptr, n, atm.log_scales.ptr,
                                                        len(atm.scales), atm.thickness, feats.ptr))
        out = ctx.alloc(n * 12)
        model.forward(feats, n, out)
        want = np.exp(out.download(np.float32, (n, 3)).astype(np.float64)).reshape(h, w, 3)
        got = img[v]
        e = np.abs(got - want) / want
        # The host restates the pixel's view cosine with numpy's float32 cos, which can differ
        # from the device's cosf by an ulp and move an FP16 input by one ulp; neighbouring
        # pixels differ by far more than such a change.
        e = e[~probe]
        print(f"{network} view {v}: render vs feature + forward entry points, median rel {np.median(e):.2e}, "
              f"99.9 % {np.quantile(e, 0.999):.2e}, max {e.max():.2e}")
        assert np.quantile(e, 0.999) < 1e-4 and e.max() < 1e-2


def test_features_match_definition(ctx):
    from skyrendering_amd import _abi
    atm = _small_atmospheres(ctx)
    q = sg.random_queries(4000, 4, atm.thickness, seed=18)
    q["mu_view"][:8] = [-1.0, 1.0, 0.0, -0.05, 0.3, -0.999, 0.999, -0.2]  # poles, horizon region
    q["mu_sun"][:4] = [0.0, -0.25, 1.0, 0.01]
    n = len(q)
    out = ctx.alloc(n * 32)
    _abi.check(_abi.lib().skyhip_surrogate_features(ctx.handle, ctx.upload(q).ptr, n, atm.log_scales.ptr,
                                                    len(atm.scales), atm.thickness, out.ptr))
    got = out.download(np.float32, (n, 8)).astype(np.float64)
    h = q["altitude"].astype(np.float64)
    mv = q["mu_view"].astype(np.float64)
    ms = q["mu_sun"].astype(np.float64)
    dp = q["delta_phi"].astype(np.float64)
    rb = 6360.0
    hh = np.maximum(h, 0.0)
    r = rb + hh
    down = np.arccos(np.sqrt(hh * (r + rb)) / r)
    up = np.pi - down
    lat = np.arccos(np.clip(mv, -1.0, 1.0))
    x_lat = np.where(lat < up, 0.5 - 0.5 * np.sqrt(np.maximum(1.0 - lat / up, 0.0)),
                     0.5 + 0.5 * np.sqrt(np.maximum((lat - up) / down, 0.0)))
    log_scales = atm.log_scales.download(np.float32, (len(atm.scales), 3)).astype(np.float64)
    want = np.stack([2.0 * np.sqrt(np.maximum(h / atm.thickness, 0.0)) - 1.0, 2.0 * x_lat - 1.0,
                     np.sign(ms) * np.sqrt(np.abs(ms)),
                     mv * ms + np.sqrt(1.0 - mv * mv) * np.sqrt(1.0 - ms * ms) * np.cos(dp),
                     log_scales[q["atmosphere"], 0] / np.log(4.0), log_scales[q["atmosphere"], 1] / np.log(4.0),
                     log_scales[q["atmosphere"], 2] / np.log(4.0), np.zeros(n)], axis=1)
    err = np.abs(got - want)
    print(f"features vs definition: max per input {err.max(axis=0)}")
    # FP32 evaluation; x_lat's square root near the horizon magnifies FP32 rounding of acos.
    assert np.quantile(err, 0.999) < 1e-5 and err.max() < 5e-3
    assert np.all(got[:, 7] == 0.0)


@pytest.mark.parametrize("network", ["bf16_silu", "f16_gate"])
def test_surrogate_gain_against_the_method_without_it(ctx, network):
    """What the trained network gains, measured: a network trained here (2^21 plain-integrator
    samples of the panoramas' domain, 200 epochs: 51,200 Adam steps, a twenty-fourth of the
    bench's training) renders 8 panoramas of 960x540 with its per-tile
    residual check and fallback, beside (a) the plain integrator it stands in for and (b) the
    method without a network: the sky-view LUT of each view (192x108, built by the batched LUT
    graph) sampled bilinearly as the reference's composite does (AtmosphereRenderer.glsl:372-379).
    Asserted, as measured on MI355X: the LUT, build included, is faster than the network and has
    the smaller median error against the plain integrator; the FP16-gate network with its
    fallback is faster than the plain integrator (1.8x here, 3.3x with the bench's full
    training; the BF16 SiLU network, ~1.0x at this training length and 1.6x fully trained, is
    reported).  The sky-view LUT is the better method on this workload; the network's gain is
    over the plain integrator only."""
    scales, views = sg.evaluation_views(8, seed=61)
    atm_train = sg.SkyAtmospheres(ctx, sg.draw_scales(512, 62))
    data = sg.SkyTrainingSet(ctx, atm_train, 1 << 21, seed=63, altitude_range=(0.0, 10.0), mu_sun_range=(0.0, 1.0))
    model = sg.SkySurrogate(ctx, seed=64, network=network)
    model.set_output_bias(data.targets_host.mean(axis=0))
    hist = model.train(data, epochs=200, batch=8192, lr=2e-3, lr_final=2e-5)
    atm = sg.SkyAtmospheres(ctx, scales, views=views)
    pano = sg.SkyPanoramas(ctx, atm, views, 960, 540)

    def timed(fn, reps=5):
        fn()
        ctx.synchronize()
        ctx.record(40)
        for _ in range(reps):
            fn()
        ctx.record(41)
        return ctx.elapsed_ms(40, 41) / reps

    plain_ms = timed(pano.render_plain)
    plain = pano.image()[..., :3].astype(np.float64)
    sur_ms = timed(lambda: pano.render_surrogate(model))
    sur = pano.image()[..., :3].astype(np.float64)
    fallback = pano.fallback_tiles() / (pano.tiles * len(views))
    lut_ms = timed(pano.render_lut)
    lut = pano.image()[..., :3].astype(np.float64)
    build_ms = timed(atm.batch.launch)
    e_sur = sg.relative_error(sur, plain, 1e-6)
    e_lut = sg.relative_error(lut, plain, 1e-6)
    print(f"{network}: final loss {hist[-1]:.3e}; plain {plain_ms:.3f} ms, network + check + fallback {sur_ms:.3f} ms "
          f"({plain_ms / sur_ms:.2f}x, {fallback:.3f} of tiles fell back), sky-view LUT {lut_ms:.3f} ms + build "
          f"{build_ms:.3f} ms; error vs plain: network median {np.median(e_sur):.2e} p99 "
          f"{np.quantile(e_sur, 0.99):.2e} max {e_sur.max():.2e}, LUT median {np.median(e_lut):.2e} p99 "
          f"{np.quantile(e_lut, 0.99):.2e} max {e_lut.max():.2e}")
    if network == "f16_gate":
        assert sur_ms < 0.85 * plain_ms
    assert lut_ms + build_ms < sur_ms
    assert np.median(e_lut) < np.median(e_sur)
