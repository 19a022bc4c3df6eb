"""Learned sky-view surrogate: a 7 -> 4 x 64 (SiLU) -> 3 MLP predicting RGB log-radiance of
the sky, trained on the GPU (skyhip_surrogate_gradient / skyhip_surrogate_adam) on samples
labelled by the plain integrator (skyhip_sky_radiance), and evaluated per pixel with a
per-tile residual check and plain-integrator fallback (skyhip_sky_render_surrogate).

Inputs: camera altitude, the view zenith (in the sky-view LUT's latitude coordinate), the
sun zenith, the sun-view cosine and the log Rayleigh / Mie / ozone scale factors of the
atmosphere (each factor in [0.25, 4] around the 100 km Earth preset).  Two networks share
the layers and parameter layout: "f16_gate" (FP16 MFMA operands, a smooth-step gated
activation in packed FP16; the default) and "bf16_silu" (BF16 MFMA operands, SiLU).  Atmospheres are built by the batched LUT builder with uniform
LUT sizes, so one SkyhipSkyLuts view covers them all.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _abi, presets
from .atmosphere import AtmosphereRenderParameters
from .batch import LutBatch
from .device import Context, DeviceBuffer

_P = C.POINTER
PARAMS = 13187
LAYERS = [(7, 64), (64, 64), (64, 64), (64, 64), (64, 3)]
SCALE_RANGE = (0.25, 4.0)
RADIANCE_FLOOR = 1e-9  # labels are log(max(L, floor))


class SkyQueryC(C.Structure):
    _fields_ = [("atmosphere", C.c_int32), ("altitude", C.c_float), ("mu_view", C.c_float), ("mu_sun", C.c_float),
                ("delta_phi", C.c_float), ("pad", C.c_float * 3)]


class SkyViewC(C.Structure):
    _fields_ = [("atmosphere", C.c_int32), ("altitude", C.c_float), ("mu_sun", C.c_float), ("pad", C.c_float)]


class SkyLutsC(C.Structure):
    _fields_ = [("atmospheres", C.c_void_p), ("transmittance", C.c_void_p), ("multiscattering", C.c_void_p),
                ("transmittance_width", C.c_int32), ("transmittance_height", C.c_int32),
                ("multiscattering_width", C.c_int32), ("multiscattering_height", C.c_int32), ("steps", C.c_float),
                ("atmosphere_count", C.c_int32)]


assert C.sizeof(SkyQueryC) == 32 and C.sizeof(SkyViewC) == 16 and C.sizeof(SkyLutsC) == 48

QUERY_DTYPE = np.dtype([("atmosphere", "<i4"), ("altitude", "<f4"), ("mu_view", "<f4"), ("mu_sun", "<f4"),
                        ("delta_phi", "<f4"), ("pad", "<f4", 3)])
VIEW_DTYPE = np.dtype([("atmosphere", "<i4"), ("altitude", "<f4"), ("mu_sun", "<f4"), ("pad", "<f4")])

_abi.register("skyhip_surrogate_parameter_count", C.c_int, [])
_abi.register("skyhip_sky_radiance", C.c_int, [C.c_void_p, _P(SkyLutsC), C.c_void_p, C.c_int32, C.c_void_p])
_abi.register("skyhip_surrogate_features", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                                     C.c_float, C.c_void_p])
_abi.register("skyhip_surrogate_forward", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                                    C.c_void_p])
_abi.register("skyhip_surrogate_workspace_bytes", C.c_size_t, [C.c_int32])
_abi.register("skyhip_surrogate_gradient", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                                     C.c_void_p, C.c_int32, C.c_float, C.c_void_p, C.c_void_p,
                                                     C.c_void_p, C.c_size_t])
_abi.register("skyhip_surrogate_adam", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32])
_abi.register("skyhip_sky_render_surrogate", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, _P(SkyLutsC), C.c_void_p,
                                                       C.c_int32, C.c_void_p, C.c_float, C.c_int32, C.c_int32,
                                                       C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p])

NETWORKS = {"f16_gate": 0, "bf16_silu": 1}  # SKYHIP_SURROGATE_F16_GATE, SKYHIP_SURROGATE_BF16_SILU
_abi.register("skyhip_sky_render_plain", C.c_int, [C.c_void_p, _P(SkyLutsC), C.c_void_p, C.c_int32, C.c_int32,
                                                   C.c_int32, C.c_void_p])
_abi.register("skyhip_sky_render_lut", C.c_int, [C.c_void_p, _P(SkyLutsC), C.c_void_p, C.c_int32, C.c_void_p,
                                                 C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p])


def init_params(seed=0) -> np.ndarray:
    """Uniform Glorot initialisation, zero biases."""
    rng = np.random.default_rng(seed)
    out = []
    for k, n in LAYERS:
        lim = math.sqrt(6.0 / (k + n))
        out.append(rng.uniform(-lim, lim, size=n * k))
        out.append(np.zeros(n))
    return np.concatenate(out).astype(np.float32)


def draw_scales(n, seed) -> np.ndarray:
    """Rayleigh, Mie and ozone scale factors, uniform in [0.25, 4]."""
    return np.random.default_rng(seed).uniform(*SCALE_RANGE, size=(n, 3))


class SkyAtmospheres:
    """Atmospheres (the 100 km Earth preset with scaled Rayleigh / Mie / ozone coefficients)
    with their transmittance and multiple-scattering LUTs built by one batched launch, and,
    optionally, sky-view LUTs for a list of views."""

    def __init__(self, ctx: Context, scales, views=None, t_size=(256, 64), ms_size=(32, 32), sky_view_size=(192, 108),
                 steps=30.0):
        self.ctx = ctx
        self.scales = np.asarray(scales, dtype=np.float64)
        base = presets.earth_100km()
        self.thickness = float(base.thickness)
        self.bottom_radius = float(base.bottom_radius)
        blocks = []
        for ray, mie, oz in self.scales:
            p = presets.earth_100km()
            p.rayleigh_scattering_scale *= ray
            p.mie_scattering_scale *= mie
            p.mie_absorption_scale *= mie
            p.ozone_absorption_scale *= oz
            blocks.append(p.buffer_data())
        self.blocks = (_abi.AtmosphereBufferData * len(blocks))(*blocks)
        self.d_blocks = ctx.alloc(C.sizeof(self.blocks))
        _abi.check(_abi.lib().skyhip_memcpy_h2d(ctx.handle, self.d_blocks.ptr, C.addressof(self.blocks),
                                                C.sizeof(self.blocks)))
        self.log_scales = ctx.upload(np.log(self.scales).astype(np.float32))
        vlist = []
        self.sky_view_size = sky_view_size
        if views is not None:
            cam = presets.bench_camera()
            vp = cam.view_projection()
            for v in views:
                theta = math.degrees(math.acos(max(-1.0, min(1.0, float(v["mu_sun"])))))
                rp = AtmosphereRenderParameters(sun_direction_theta=theta, sun_direction_phi=180.0,
                                                sky_view_lut_steps=steps, camera_position=(0.0, float(v["altitude"]), 0.0),
                                                view_projection=vp)
                vlist.append((int(v["atmosphere"]), rp.buffer_data(self.bottom_radius), sky_view_size))
        self.batch = LutBatch(ctx, [(b, t_size, ms_size) for b in blocks], vlist)
        self.batch.launch()
        ctx.synchronize()
        t0, m0 = self.batch.set_output_ptrs(0)
        self.luts = SkyLutsC(atmospheres=self.d_blocks.ptr, transmittance=t0, multiscattering=m0,
                             transmittance_width=t_size[0], transmittance_height=t_size[1],
                             multiscattering_width=ms_size[0], multiscattering_height=ms_size[1], steps=steps,
                             atmosphere_count=len(blocks))
        if vlist:
            ptrs = np.array([self.batch.view_output_ptrs(i)[0] for i in range(len(vlist))], dtype=np.uint64)
            self.sky_view_ptrs = ctx.upload(ptrs)

    def radiance(self, queries: DeviceBuffer, n, out: DeviceBuffer):
        _abi.check(_abi.lib().skyhip_sky_radiance(self.ctx.handle, C.byref(self.luts), queries.ptr, n, out.ptr))

    def host_luts(self):
        """(blocks, T [a, th, tw, 4], MS [a, mh, mw, 4]) host copies for the oracle."""
        T = np.stack([self.batch.set_outputs(s)[0] for s in range(len(self.scales))])
        MS = np.stack([self.batch.set_outputs(s)[1] for s in range(len(self.scales))])
        return self.blocks, T, MS


def brightness_weights(radiance, reference=1e-2) -> np.ndarray:
    """Loss weight per sample: min(1, Y / reference)^(1/2), Y the mean RGB radiance.  Samples
    stay uniform over the input space; a dark sky (night, looking up from the top of the
    atmosphere) counts less, so the network's capacity goes where the sky is seen."""
    y = np.asarray(radiance, dtype=np.float64).mean(axis=1)
    return np.sqrt(np.minimum(1.0, np.maximum(y, 0.0) / reference)).astype(np.float32)


def random_queries(n, n_atm, thickness, seed, altitude_range=None, mu_sun_range=(-1.0, 1.0)) -> np.ndarray:
    """Uniform samples: atmosphere, altitude in altitude_range (default [0, thickness]), mu_view
    in [-1, 1], mu_sun in mu_sun_range, delta_phi in [0, pi]."""
    rng = np.random.default_rng(seed)
    lo, hi = altitude_range if altitude_range is not None else (0.0, thickness)
    q = np.zeros(n, QUERY_DTYPE)
    q["atmosphere"] = rng.integers(0, n_atm, n)
    q["altitude"] = rng.uniform(lo, hi, n)
    q["mu_view"] = rng.uniform(-1.0, 1.0, n)
    q["mu_sun"] = rng.uniform(*mu_sun_range, n)
    q["delta_phi"] = rng.uniform(0.0, math.pi, n)
    return q


class SkyTrainingSet:
    """Queries, surrogate inputs and plain-integrator labels (log radiance), on the device."""

    def __init__(self, ctx: Context, atmospheres: SkyAtmospheres, n, seed=21, altitude_range=None,
                 mu_sun_range=(-1.0, 1.0)):
        self.ctx, self.n = ctx, n
        self.atmospheres = atmospheres
        self.queries_host = random_queries(n, len(atmospheres.scales), atmospheres.thickness, seed, altitude_range,
                                           mu_sun_range)
        self.queries = ctx.upload(self.queries_host)
        self.features = ctx.alloc(n * 32)
        _abi.check(_abi.lib().skyhip_surrogate_features(ctx.handle, self.queries.ptr, n, atmospheres.log_scales.ptr,
                                                        len(atmospheres.scales), atmospheres.thickness,
                                                        self.features.ptr))
        rad = ctx.alloc(n * 12)
        atmospheres.radiance(self.queries, n, rad)
        self.radiance = rad.download(np.float32, (n, 3))
        rad.free()
        self.targets_host = np.log(np.maximum(self.radiance, RADIANCE_FLOOR)).astype(np.float32)
        self.targets = ctx.upload(self.targets_host)
        self.weights_host = brightness_weights(self.radiance)
        self.weights = ctx.upload(self.weights_host)


class SkySurrogate:
    """The network's parameters and Adam state on the device."""

    def __init__(self, ctx: Context, params=None, seed=0, network="f16_gate"):
        assert _abi.lib().skyhip_surrogate_parameter_count() == PARAMS
        if network not in NETWORKS:
            raise ValueError(f"network must be one of {sorted(NETWORKS)}")
        self.ctx = ctx
        self.network = network
        self.network_id = NETWORKS[network]
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.params = ctx.upload(p)
        self.m = ctx.alloc(PARAMS * 4).zero()
        self.v = ctx.alloc(PARAMS * 4).zero()
        self.grad = ctx.alloc(PARAMS * 4)
        self.loss = ctx.alloc(4)
        self.step = 0
        self._ws = None
        self._ws_n = 0

    def host_params(self):
        return self.params.download(np.float32, (PARAMS,))

    def set_output_bias(self, mean_log_radiance):
        p = self.host_params()
        p[13184:13187] = mean_log_radiance
        self.params.upload(p)

    def forward(self, features: DeviceBuffer, n, out: DeviceBuffer):
        _abi.check(_abi.lib().skyhip_surrogate_forward(self.ctx.handle, self.network_id, self.params.ptr, features.ptr,
                                                       n, out.ptr))

    def gradient(self, features_ptr, targets_ptr, n, loss_scale, weights_ptr=None):
        need = _abi.lib().skyhip_surrogate_workspace_bytes(n)
        if self._ws is None or self._ws_n < need:
            if self._ws is not None:
                self._ws.free()
            self._ws = self.ctx.alloc(need)
            self._ws_n = need
        _abi.check(_abi.lib().skyhip_surrogate_gradient(self.ctx.handle, self.network_id, self.params.ptr, features_ptr,
                                                        targets_ptr, weights_ptr, n, loss_scale, self.grad.ptr,
                                                        self.loss.ptr, self._ws.ptr, self._ws_n))

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _abi.check(_abi.lib().skyhip_surrogate_adam(self.ctx.handle, self.params.ptr, self.grad.ptr, self.m.ptr,
                                                    self.v.ptr, lr, beta1, beta2, eps, self.step))

    def train(self, data: SkyTrainingSet, epochs, batch=8192, lr=3e-3, lr_final=3e-5, log_every=0):
        """Adam with a cosine learning-rate decay over `epochs` passes in fixed batch order;
        returns the mean loss (brightness-weighted mean squared log-radiance error) per epoch."""
        steps_per_epoch = data.n // batch
        total = steps_per_epoch * epochs
        history = []
        k = 0
        for e in range(epochs):
            losses = []
            for b in range(steps_per_epoch):
                rate = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * k / max(total, 1)))
                self.gradient(data.features.ptr + b * batch * 32, data.targets.ptr + b * batch * 12, batch,
                              1.0 / (3.0 * batch), data.weights.ptr + b * batch * 4)
                self.adam(rate)
                k += 1
                if b % 64 == 0:
                    losses.append(float(self.loss.download(np.float32, (1,))[0]))
            history.append(float(np.mean(losses)))
            if log_every and (e % log_every == 0 or e == epochs - 1):
                print(f"epoch {e}: loss {history[-1]:.4e}", flush=True)
        return history


class SkyPanoramas:
    """Equirectangular sky panoramas over a set of views (atmosphere, camera altitude, sun),
    rendered three ways: plain integrator per pixel, surrogate with residual-check fallback,
    and the sky-view LUT sampled bilinearly."""

    def __init__(self, ctx: Context, atmospheres: SkyAtmospheres, views, width, height):
        self.ctx, self.atm = ctx, atmospheres
        self.views_host = np.asarray(views, VIEW_DTYPE)
        self.views = ctx.upload(self.views_host)
        self.nv = len(self.views_host)
        self.width, self.height = width, height
        self.tiles = ((width + 15) // 16) * ((height + 15) // 16)
        self.npix = width * height * self.nv
        self.out = ctx.alloc(self.npix * 16)
        self.flags = ctx.alloc(self.tiles * self.nv * 4)
        self.counter = ctx.alloc(4)

    def render_plain(self):
        _abi.check(_abi.lib().skyhip_sky_render_plain(self.ctx.handle, C.byref(self.atm.luts), self.views.ptr,
                                                      self.nv, self.width, self.height, self.out.ptr))

    def render_surrogate(self, model: SkySurrogate, tolerance=1e-2, radiance_floor=1e-6):
        _abi.check(_abi.lib().skyhip_sky_render_surrogate(
            self.ctx.handle, model.network_id, model.params.ptr, C.byref(self.atm.luts), self.views.ptr, self.nv,
            self.atm.log_scales.ptr, self.atm.thickness, self.width, self.height, tolerance, radiance_floor,
            self.out.ptr, self.flags.ptr, self.counter.ptr))

    def render_lut(self):
        sw, sh = self.atm.sky_view_size
        _abi.check(_abi.lib().skyhip_sky_render_lut(self.ctx.handle, C.byref(self.atm.luts), self.views.ptr,
                                                    self.nv, self.atm.sky_view_ptrs.ptr, sw, sh, self.width,
                                                    self.height, self.out.ptr))

    def fallback_tiles(self) -> int:
        return int(self.counter.download(np.uint32, (1,))[0])

    def image(self):
        return self.out.download(np.float32, (self.nv, self.height, self.width, 4))

    def pixel_queries(self, view, xs, ys) -> np.ndarray:
        """The queries of pixels (xs, ys) of a view, as [n, 5] doubles for the oracle."""
        v = self.views_host[view]
        mu_v = np.cos(np.pi * (np.asarray(ys, np.float32) + np.float32(0.5)) / np.float32(self.height))
        dphi = np.abs(np.float32(2 * np.pi) * (np.asarray(xs, np.float32) + np.float32(0.5)) /
                      np.float32(self.width) - np.float32(np.pi))
        return np.stack([np.full(len(xs), v["atmosphere"], np.float64), np.full(len(xs), v["altitude"], np.float64),
                         mu_v.astype(np.float64), np.full(len(xs), v["mu_sun"], np.float64),
                         dphi.astype(np.float64)], axis=1)


def evaluation_views(n=64, seed=31, max_altitude=10.0, max_sun_zenith_deg=90.0):
    """(scales [n, 3], views): one atmosphere per view, camera altitude U[0, max_altitude] km,
    sun zenith U[0, max_sun_zenith_deg]."""
    rng = np.random.default_rng(seed)
    scales = draw_scales(n, seed + 1)
    v = np.zeros(n, VIEW_DTYPE)
    v["atmosphere"] = np.arange(n)
    v["altitude"] = rng.uniform(0.0, max_altitude, n)
    v["mu_sun"] = np.cos(np.radians(rng.uniform(0.0, max_sun_zenith_deg, n)))
    return scales, v


def relative_error(x, ref, floor=1e-6):
    """max over RGB of |x - ref| / max(ref, floor), per pixel."""
    return np.max(np.abs(x - ref) / np.maximum(ref, floor), axis=-1)
