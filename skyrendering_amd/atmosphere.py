"""Host mirror of the reference's atmosphere classes over the skyhip C ABI.

Names, fields, defaults and units follow src/SkyRendering/Atmosphere.h:10-92,
AtmosphereRenderer.h:12-132, Camera.h and Earth.h, so a caller of the reference's
Atmosphere::UpdateLuts / AtmosphereRenderer::Render finds the same surface here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields

import numpy as np

from . import _abi
from .device import Context, DeviceBuffer


def _vec(ctype_arr, values):
    for i, v in enumerate(values):
        ctype_arr[i] = float(v)


@dataclass
class AtmosphereParameters:
    """AtmosphereParameters, src/SkyRendering/Atmosphere.h:10-38 (degrees, km, per km)."""
    solar_illuminance: tuple = (1.0, 1.0, 1.0)
    sun_angular_radius: float = 0.5334 * 0.5
    bottom_radius: float = 6360.0
    thickness: float = 6420.0 - 6360.0
    ground_albedo: tuple = (0.4, 0.4, 0.4)
    rayleigh_exponential_distribution: float = 8.0
    rayleigh_scattering_scale: float = 0.0331
    rayleigh_scattering: tuple = (0.175287, 0.409607, 1.0)
    mie_exponential_distribution: float = 1.2
    mie_phase_g: float = 0.8
    mie_scattering_scale: float = 0.003996
    mie_scattering: tuple = (1.0, 1.0, 1.0)
    mie_absorption_scale: float = 0.000444
    mie_absorption: tuple = (1.0, 1.0, 1.0)
    ozone_center_altitude: float = 25.0
    ozone_width: float = 15.0
    ozone_absorption_scale: float = 0.001881
    ozone_absorption: tuple = (0.345561, 1.0, 0.045189)
    transmittance_steps: float = 40.0
    multiscattering_steps: float = 30.0
    multiscattering_mask: float = 1.0

    def to_c(self) -> _abi.AtmosphereParametersC:
        c = _abi.AtmosphereParametersC()
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, (tuple, list)):
                _vec(getattr(c, f.name), v)
            else:
                setattr(c, f.name, float(v))
        return c

    def buffer_data(self) -> _abi.AtmosphereBufferData:
        """AssignBufferData (Atmosphere.cpp:49-70) through the C ABI."""
        out = _abi.AtmosphereBufferData()
        _abi.check(_abi.lib().skyhip_atmosphere_assign_buffer_data(C.byref(self.to_c()), C.byref(out)))
        return out

    @classmethod
    def from_json(cls, d: dict) -> "AtmosphereParameters":
        """Reads the `earth_.parameters` object of bin/config*.json (missing keys keep defaults)."""
        p = cls()
        for f in fields(cls):
            if f.name in d:
                v = d[f.name]
                setattr(p, f.name, tuple(v) if isinstance(v, list) else float(v))
        return p


@dataclass
class Camera:
    """Camera, src/Base/include/Camera.h (defaults from Camera.h:38-40, 65)."""
    position: tuple = (0.0, 0.0, 0.0)
    yaw: float = 0.0
    pitch: float = 0.0
    aspect: float = 1.0
    fovy: float = 45.0
    zNear: float = 1e-1
    zFar: float = 1e3

    def to_c(self) -> _abi.CameraC:
        c = _abi.CameraC()
        _vec(c.position, self.position)
        c.yaw, c.pitch, c.aspect, c.fovy, c.z_near, c.z_far = (
            self.yaw, self.pitch, self.aspect, self.fovy, self.zNear, self.zFar)
        return c

    def view_projection(self) -> np.ndarray:
        """ProjectionMatrix() * ViewMatrix() as column-major float32[16]."""
        vp = (C.c_float * 16)()
        front, right, up = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
        _abi.check(_abi.lib().skyhip_camera_view_projection(C.byref(self.to_c()), vp, front, right, up))
        return np.array(vp[:], dtype=np.float32)

    def basis(self):
        vp = (C.c_float * 16)()
        front, right, up = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
        _abi.check(_abi.lib().skyhip_camera_view_projection(C.byref(self.to_c()), vp, front, right, up))
        return np.array(front[:]), np.array(right[:]), np.array(up[:])


@dataclass
class AtmosphereRenderParameters:
    """AtmosphereRenderParameters, src/SkyRendering/AtmosphereRenderer.h:12-35."""
    sun_direction_theta: float = 70.0
    sun_direction_phi: float = 180.0
    star_luminance_scale: float = 0.005
    raymarching_steps: float = 40.0
    sky_view_lut_steps: float = 40.0
    aerial_perspective_lut_steps: float = 40.0
    aerial_perspective_lut_max_distance: float = 100.0
    camera_position: tuple = (0.0, 0.0, 0.0)
    view_projection: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32).reshape(-1))

    def to_c(self) -> _abi.AtmosphereRenderParametersC:
        c = _abi.AtmosphereRenderParametersC()
        c.sun_direction_theta = self.sun_direction_theta
        c.sun_direction_phi = self.sun_direction_phi
        c.star_luminance_scale = self.star_luminance_scale
        c.raymarching_steps = self.raymarching_steps
        c.sky_view_lut_steps = self.sky_view_lut_steps
        c.aerial_perspective_lut_steps = self.aerial_perspective_lut_steps
        c.aerial_perspective_lut_max_distance = self.aerial_perspective_lut_max_distance
        _vec(c.camera_position, self.camera_position)
        _vec(c.view_projection, np.asarray(self.view_projection, dtype=np.float32).reshape(-1))
        return c

    def buffer_data(self, bottom_radius: float) -> _abi.AtmosphereRenderBufferData:
        """AssignBufferData (AtmosphereRenderer.cpp:52-83) through the C ABI."""
        out = _abi.AtmosphereRenderBufferData()
        _abi.check(_abi.lib().skyhip_atmosphere_render_assign_buffer_data(
            C.byref(self.to_c()), float(bottom_radius), C.byref(out)))
        return out


@dataclass
class AtmosphereRenderInitParameters:
    """AtmosphereRenderInitParameters, AtmosphereRenderer.h:44-66, plus the LUT sizes the
    reference fixes at compile time (AtmosphereRenderer.cpp:15-21)."""
    sky_view_lut_dither_sample_point_enable: bool = False
    aerial_perspective_lut_dither_sample_point_enable: bool = False
    aerial_perspective_lut_depth: int = 32
    sky_view_lut_size: tuple = (128, 128)
    aerial_perspective_lut_size: tuple = (32, 32)


class Atmosphere:
    """Atmosphere (src/SkyRendering/Atmosphere.h:70-92): owns the T (256x64) and MS (32x32)
    RGBA32F LUTs and rebuilds them in UpdateLuts."""

    def __init__(self, ctx: Context, transmittance_size=(256, 64), multiscattering_size=(32, 32)):
        self.ctx = ctx
        self.transmittance_size = tuple(transmittance_size)
        self.multiscattering_size = tuple(multiscattering_size)
        tw, th = self.transmittance_size
        mw, mh = self.multiscattering_size
        self.transmittance_texture = ctx.alloc(tw * th * 16)
        self.multiscattering_texture = ctx.alloc(mw * mh * 16)
        self.buffer = None

    def update_luts(self, parameters: AtmosphereParameters):
        """Atmosphere::UpdateLuts (Atmosphere.cpp:101-124)."""
        tw, th = self.transmittance_size
        mw, mh = self.multiscattering_size
        self.buffer = parameters.buffer_data()
        _abi.check(_abi.lib().skyhip_atmosphere_update_luts(
            self.ctx.handle, C.byref(parameters.to_c()), self.transmittance_texture.ptr, tw, th,
            self.multiscattering_texture.ptr, mw, mh))

    UpdateLuts = update_luts

    def transmittance(self) -> np.ndarray:
        tw, th = self.transmittance_size
        return self.transmittance_texture.download(np.float32, (th, tw, 4))

    def multiscattering(self) -> np.ndarray:
        mw, mh = self.multiscattering_size
        return self.multiscattering_texture.download(np.float32, (mh, mw, 4))


class AtmosphereRenderer:
    """AtmosphereRenderer's LUT passes (AtmosphereRenderer.cpp:164-233): sky view and
    aerial perspective.  The environment cube and IBL that consume the sky-view LUT are
    ibl.EnvironmentLuminance and ibl.IBL; the fullscreen composite is not part of this
    library."""

    def __init__(self, ctx: Context, init: AtmosphereRenderInitParameters | None = None,
                 blue_noise: DeviceBuffer | None = None, blue_noise_size: int = 64):
        self.ctx = ctx
        self.init = init or AtmosphereRenderInitParameters()
        sw, sh = self.init.sky_view_lut_size
        aw, ah = self.init.aerial_perspective_lut_size
        ad = self.init.aerial_perspective_lut_depth
        self.sky_view_luminance_texture = ctx.alloc(sw * sh * 16)
        self.sky_view_transmittance_texture = ctx.alloc(sw * sh * 16)
        self.aerial_perspective_luminance_texture = ctx.alloc(aw * ah * ad * 16)
        self.aerial_perspective_transmittance_texture = ctx.alloc(aw * ah * ad * 16)
        self.blue_noise = blue_noise
        self.blue_noise_size = blue_noise_size
        self.buffer = None
        self.sun_direction = None

    def render(self, atmosphere: Atmosphere, earth: AtmosphereParameters, parameters: AtmosphereRenderParameters,
               sky_view=True, aerial_perspective=True):
        """AtmosphereRenderer::Render, LUT part (AtmosphereRenderer.cpp:164-233)."""
        L = _abi.lib()
        self.buffer = parameters.buffer_data(earth.bottom_radius)
        self.sun_direction = np.array(self.buffer.sun_direction[:], dtype=np.float32)
        self.aerial_perspective_lut_max_distance = parameters.aerial_perspective_lut_max_distance
        tw, th = atmosphere.transmittance_size
        mw, mh = atmosphere.multiscattering_size
        sw, sh = self.init.sky_view_lut_size
        aw, ah = self.init.aerial_perspective_lut_size
        ad = self.init.aerial_perspective_lut_depth
        bn = self.blue_noise.ptr if self.blue_noise is not None else None
        if sky_view:
            _abi.check(L.skyhip_sky_view_lut(
                self.ctx.handle, C.byref(atmosphere.buffer), C.byref(self.buffer),
                atmosphere.transmittance_texture.ptr, tw, th, atmosphere.multiscattering_texture.ptr, mw, mh,
                bn if self.init.sky_view_lut_dither_sample_point_enable else None, self.blue_noise_size,
                self.sky_view_luminance_texture.ptr, self.sky_view_transmittance_texture.ptr, sw, sh))
        if aerial_perspective:
            _abi.check(L.skyhip_aerial_perspective_lut(
                self.ctx.handle, C.byref(atmosphere.buffer), C.byref(self.buffer),
                atmosphere.transmittance_texture.ptr, tw, th, atmosphere.multiscattering_texture.ptr, mw, mh,
                bn if self.init.aerial_perspective_lut_dither_sample_point_enable else None, self.blue_noise_size,
                self.aerial_perspective_luminance_texture.ptr, self.aerial_perspective_transmittance_texture.ptr,
                aw, ah, ad))

    Render = render

    def sky_view(self):
        sw, sh = self.init.sky_view_lut_size
        return (self.sky_view_luminance_texture.download(np.float32, (sh, sw, 4)),
                self.sky_view_transmittance_texture.download(np.float32, (sh, sw, 4)))

    def aerial_perspective(self):
        aw, ah = self.init.aerial_perspective_lut_size
        ad = self.init.aerial_perspective_lut_depth
        return (self.aerial_perspective_luminance_texture.download(np.float32, (ad, ah, aw, 4)),
                self.aerial_perspective_transmittance_texture.download(np.float32, (ad, ah, aw, 4)))
