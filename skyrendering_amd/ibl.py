"""The atmosphere renderer's environment-luminance cube and the image-based-lighting
precomputation that consumes it (AtmosphereRenderer::Render's environment pass,
src/SkyRendering/AtmosphereRenderer.cpp:234-244, and IBL, src/Base/src/IBL.cpp)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .atmosphere import Atmosphere, AtmosphereRenderer
from .device import Context, DeviceBuffer

_abi.register("skyhip_cube_texels", C.c_longlong, [C.c_int32, C.c_int32])
_abi.register("skyhip_environment_luminance", C.c_int, [
    C.c_void_p, C.POINTER(_abi.AtmosphereBufferData), C.POINTER(_abi.AtmosphereRenderBufferData), C.c_void_p,
    C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32])
_abi.register("skyhip_ibl_precompute", C.c_int, [
    C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32])

ENVIRONMENT_LUMINANCE_WIDTH = 128  # kEnvironmentLuminanceTextureWidth, AtmosphereRenderer.cpp:23


def mip_levels(width: int) -> int:
    return int(width).bit_length()


def cube_texels(width: int, levels: int) -> int:
    return int(_abi.lib().skyhip_cube_texels(width, levels))


def split_cube(flat: np.ndarray, width: int, levels: int) -> list[np.ndarray]:
    """Per-level views [6, w_l, w_l, 4] of a flat cube array (level-major)."""
    out, off = [], 0
    for lvl in range(levels):
        w = max(width >> lvl, 1)
        n = 6 * w * w
        out.append(flat[off:off + n].reshape(6, w, w, 4))
        off += n
    return out


class EnvironmentLuminance:
    """The environment-luminance cube (RGBA16F, full mip chain) of an atmosphere renderer."""

    def __init__(self, ctx: Context, width: int = ENVIRONMENT_LUMINANCE_WIDTH):
        self.ctx, self.width = ctx, width
        self.levels = mip_levels(width)
        self.texture = ctx.alloc(cube_texels(width, self.levels) * 8)

    def update(self, atmosphere: Atmosphere, renderer: AtmosphereRenderer):
        """Recomputes the cube from the renderer's current sky-view LUT pair and the
        atmosphere's transmittance LUT (after AtmosphereRenderer.render)."""
        tw, th = atmosphere.transmittance_size
        sw, sh = renderer.init.sky_view_lut_size
        _abi.check(_abi.lib().skyhip_environment_luminance(
            self.ctx.handle, C.byref(atmosphere.buffer), C.byref(renderer.buffer),
            atmosphere.transmittance_texture.ptr, tw, th, renderer.sky_view_luminance_texture.ptr,
            renderer.sky_view_transmittance_texture.ptr, sw, sh, self.texture.ptr, self.width))

    def download(self) -> list[np.ndarray]:
        """float16 [6, w_l, w_l, 4] per level."""
        flat = self.texture.download(np.float16, (cube_texels(self.width, self.levels), 4))
        return split_cube(flat, self.width, self.levels)


class IBL:
    """IBL (src/Base/include/IBL.h): SH9 radiance coefficients and the prefiltered
    radiance cube of an environment cube."""

    kPrefilteredRadianceResolution = 128
    kRoughnessCount = 5

    def __init__(self, ctx: Context, resolution: int = kPrefilteredRadianceResolution,
                 roughness_count: int = kRoughnessCount):
        self.ctx, self.resolution, self.roughness_count = ctx, resolution, roughness_count
        self.env_radiance_sh_buffer = ctx.alloc(9 * 16)
        self.prefiltered_radiance_texture = ctx.alloc(cube_texels(resolution, roughness_count) * 8)

    def precompute(self, environment: EnvironmentLuminance):
        """IBL::Precompute."""
        _abi.check(_abi.lib().skyhip_ibl_precompute(
            self.ctx.handle, environment.texture.ptr, environment.width, self.env_radiance_sh_buffer.ptr,
            self.prefiltered_radiance_texture.ptr, self.resolution, self.roughness_count))

    Precompute = precompute

    def env_radiance_sh(self) -> np.ndarray:
        """[9, 4] float32 (a = 0)."""
        return self.env_radiance_sh_buffer.download(np.float32, (9, 4))

    def prefiltered_radiance(self) -> list[np.ndarray]:
        flat = self.prefiltered_radiance_texture.download(
            np.float16, (cube_texels(self.resolution, self.roughness_count), 4))
        return split_cube(flat, self.resolution, self.roughness_count)


def upload_cube(ctx: Context, levels: list[np.ndarray]) -> DeviceBuffer:
    """A device cube from per-level float16 arrays [6, w, w, 4]."""
    return ctx.upload(np.ascontiguousarray(np.concatenate([np.asarray(l, np.float16).reshape(-1, 4) for l in levels])))
