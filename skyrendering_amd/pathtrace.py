"""Voxel-cloud path tracing: VolumetricCloud::PathTracing (src/SkyRendering/VolumetricCloud.h:140-190,
VolumetricCloud.cpp:495-581) over the HIP kernels of skyhip_cloud_path_trace.

`PathTracing` keeps the reference's object model: an RGBA32F accumulation image and an R8
rendered mask at the viewport size, a frame counter that starts at 1, a tile schedule of
sqrt_tile_count^2 regions (GetRenderRegion), and a display pass that composites the running
mean over an RGBA16F image.  `VoxelGrid` builds the FP16 density volume of the path-tracing
workload and its 16^3-brick majorants on the device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi, presets
from .atmosphere import (Atmosphere, AtmosphereRenderer, AtmosphereRenderInitParameters,
                         AtmosphereRenderParameters, Camera)
from .device import Context, DeviceBuffer

_P = C.POINTER
_abi.register("skyhip_voxel_grid_generate", C.c_int, [C.c_void_p, _P(_abi.VoxelGridCreateInfoC), C.c_void_p,
                                                      C.c_int32, C.c_int32, C.c_int32])
_abi.register("skyhip_voxel_grid_brick_size", C.c_int, [])
_abi.register("skyhip_voxel_grid_majorants_sized", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                                            C.c_int32, C.c_void_p])
_abi.register("skyhip_voxel_grid_majorants", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                                       C.c_void_p])
_abi.register("skyhip_voxel_grid_cells", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                                   C.c_void_p])
_abi.register("skyhip_cloud_path_trace", C.c_int, [
    C.c_void_p, _P(_abi.AtmosphereBufferData), _P(_abi.AtmosphereRenderBufferData), _P(_abi.CloudCommonBufferData),
    _P(_abi.CloudMaterial), _P(_abi.PathTracingParametersC), _P(_abi.PathTracingInputsC), _P(C.c_int32),
    C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p])
_abi.register("skyhip_cloud_path_trace_display", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                           C.c_void_p, C.c_int32, C.c_int32])

PRNG = {"WangHash": _abi.PRNG_WANG_HASH, "PCGHash": _abi.PRNG_PCG_HASH}
ENVIRONMENT_LIGHTING = {"OFF": _abi.ENV_OFF, "CONST_ENVIRONMENT_MAP": _abi.ENV_CONST_ENVIRONMENT_MAP,
                        "GROUND_SINGLE_BOUNCE": _abi.ENV_GROUND_SINGLE_BOUNCE,
                        "GROUND_MULTI_BOUNCE": _abi.ENV_GROUND_MULTI_BOUNCE}


@dataclass
class PathTracingInitParam:
    """VolumetricCloud::PathTracing::InitParam (VolumetricCloud.h:158-168), plus the
    single-scattering albedo (the reference's is 1)."""
    sqrt_tile_count: int = 1
    max_bounces: int = 128
    region_box_half_width: float = 100.0
    importance_sampling: bool = True
    forward_phase_g: float = 0.85
    back_phase_g: float = -0.15
    forward_scattering_ratio: float = 0.7
    prng: str = "PCGHash"
    environment_lighting: str = "GROUND_MULTI_BOUNCE"
    albedo: float = 1.0

    def to_c(self, sigma_t_max: float, model) -> _abi.PathTracingParametersC:
        """The shader constants VolumetricCloud.cpp:506-519 bakes in: kSigmaTMax is the
        material's GetSigmaTMax(), kModelMatrix3 the upper 3x3 of the cloud's model matrix."""
        if self.prng not in PRNG:
            raise ValueError(f"unknown PRNG {self.prng!r}")
        if self.environment_lighting not in ENVIRONMENT_LIGHTING:
            raise ValueError(f"unknown environment lighting {self.environment_lighting!r}")
        c = _abi.PathTracingParametersC()
        c.max_bounces = int(self.max_bounces)
        c.region_box_half_width = float(self.region_box_half_width)
        c.importance_sampling = 1 if self.importance_sampling else 0
        c.forward_phase_g = float(self.forward_phase_g)
        c.back_phase_g = float(self.back_phase_g)
        c.forward_scattering_ratio = float(self.forward_scattering_ratio)
        c.prng = PRNG[self.prng]
        c.environment_lighting = ENVIRONMENT_LIGHTING[self.environment_lighting]
        c.sigma_t_max = float(sigma_t_max)
        c.albedo = float(self.albedo)
        m = np.asarray(model, dtype=np.float32).reshape(4, 4)  # column-major 4x4: m[col][row]
        for col in range(3):
            for row in range(3):
                c.model3[col * 3 + row] = float(m[col][row])
        return c


@dataclass
class VoxelGridCreateInfo:
    """The path-tracing workload's density volume (see skyhip_voxel_grid_generate)."""
    seed: int = 8
    base_frequency: int = 4
    octaves: int = 5
    fbm_inv_std: float = 1.0 / 0.162  # 5-octave fBm standard deviation measured over a 96^3 sample
    lognormal_sigma: float = 1.2
    threshold: float = 1.0
    ellipsoid_center: tuple = (0.5, 0.5, 0.45)
    ellipsoid_radii: tuple = (0.45, 0.40, 0.45)

    def to_c(self) -> _abi.VoxelGridCreateInfoC:
        c = _abi.VoxelGridCreateInfoC(seed=self.seed, base_frequency=self.base_frequency, octaves=self.octaves,
                                      fbm_inv_std=self.fbm_inv_std, lognormal_sigma=self.lognormal_sigma,
                                      threshold=self.threshold)
        for k in range(3):
            c.ellipsoid_center[k] = self.ellipsoid_center[k]
            c.ellipsoid_radii[k] = self.ellipsoid_radii[k]
        return c


class VoxelGrid:
    """FP16 density grid (x fastest), its brick majorants (brick edge `brick` texels, the
    library's default when None) and, with `cells`, its cell-corner copy (eight texels per
    cell, one 16-byte load per trilinear lookup; 4.2x the grid's memory), resident on the
    device."""

    def __init__(self, ctx: Context, dims=(512, 512, 512), brick=None, cells=True):
        self.ctx = ctx
        self.dims = tuple(int(d) for d in dims)
        dx, dy, dz = self.dims
        self.density = ctx.alloc(dx * dy * dz * 2)
        b = int(brick) if brick else _abi.lib().skyhip_voxel_grid_brick_size()
        self.brick = b
        self.brick_dims = tuple((d + b - 1) // b for d in self.dims)
        bx, by, bz = self.brick_dims
        self.majorants = ctx.alloc(bx * by * bz * 4)
        self.cells = ctx.alloc((dx + 1) * (dy + 1) * (dz + 1) * 16) if cells else None

    def generate(self, info: VoxelGridCreateInfo | None = None):
        info = info or VoxelGridCreateInfo()
        _abi.check(_abi.lib().skyhip_voxel_grid_generate(self.ctx.handle, C.byref(info.to_c()), self.density.ptr,
                                                         *self.dims))
        self.build_majorants()
        return self

    def upload(self, grid_f16_bits: np.ndarray):
        g = np.ascontiguousarray(grid_f16_bits, dtype=np.uint16)
        if g.shape != self.dims[::-1]:
            raise ValueError(f"grid shape {g.shape} != {self.dims[::-1]}")
        self.density.upload(g)
        self.build_majorants()
        return self

    def build_majorants(self):
        """The grid's derived data: brick majorants and (when kept) the cell-corner copy."""
        _abi.check(_abi.lib().skyhip_voxel_grid_majorants_sized(self.ctx.handle, self.density.ptr, *self.dims,
                                                                self.brick, self.majorants.ptr))
        if self.cells is not None:
            _abi.check(_abi.lib().skyhip_voxel_grid_cells(self.ctx.handle, self.density.ptr, *self.dims,
                                                          self.cells.ptr))

    def download_cells(self):
        """Cell-corner copy as uint16[dz + 1, dy + 1, dx + 1, 8] (cell (-1, -1, -1) first)."""
        dx, dy, dz = self.dims
        return self.cells.download(np.uint16, (dz + 1, dy + 1, dx + 1, 8))

    def download(self):
        dx, dy, dz = self.dims
        return self.density.download(np.uint16, (dz, dy, dx))

    def download_majorants(self):
        bx, by, bz = self.brick_dims
        return self.majorants.download(np.float32, (bz, by, bx))

    def max_density(self) -> float:
        return float(self.download_majorants().max())


def voxel_f16_material(grid: VoxelGrid, frequency, bias, density) -> _abi.CloudMaterial:
    """A VOXEL_F16 material: uv = pos.xy * frequency + bias (the voxel material's mapping,
    VolumetricCloudMaterialVoxel.glsl:12-16), w = height01, sigma = grid * density."""
    m = _abi.CloudMaterial()
    m.kind = _abi.MATERIAL_VOXEL_F16
    m.voxel.uSampleFrequency[0], m.voxel.uSampleFrequency[1] = frequency
    m.voxel.uSampleBias[0], m.voxel.uSampleBias[1] = bias
    m.voxel.uDensity = density
    m.voxel_grid = grid.density.ptr
    for k in range(3):
        m.voxel_grid_dim[k] = grid.dims[k]
    return m


def path_trace(ctx: Context, atmosphere_block, render_block, common, material, params, inputs, region, first_frame,
               frames, accumulation: DeviceBuffer, mask: DeviceBuffer | None, width, height,
               work_counters: DeviceBuffer | None = None):
    reg = (C.c_int32 * 4)(*[int(x) for x in region])
    _abi.check(_abi.lib().skyhip_cloud_path_trace(
        ctx.handle, C.byref(atmosphere_block), C.byref(render_block), C.byref(common), C.byref(material),
        C.byref(params), C.byref(inputs), reg, int(first_frame), int(frames), accumulation.ptr,
        mask.ptr if mask is not None else None, width, height,
        work_counters.ptr if work_counters is not None else None))


class PathTracing:
    """VolumetricCloud::PathTracing: accumulation, tile schedule and display."""

    def __init__(self, ctx: Context, viewport, init: PathTracingInitParam, atmosphere_block, render_block, common,
                 material, inputs: _abi.PathTracingInputsC, sigma_t_max: float, model):
        self.ctx = ctx
        self.width, self.height = viewport
        self.init = init
        self.sqrt_tile_count = init.sqrt_tile_count
        self.atmosphere_block, self.render_block, self.common = atmosphere_block, render_block, common
        self.material, self.inputs = material, inputs
        self.params = init.to_c(sigma_t_max, model)
        n = self.width * self.height
        self.accumulation = ctx.alloc(n * 16)
        self.accumulation.zero()
        self.rendered_mask = ctx.alloc(n)
        self.frame_cnt = 0
        self.tile_index = 0

    def get_render_region(self):
        """GetRenderRegion (VolumetricCloud.cpp:568-581)."""
        s = self.sqrt_tile_count
        x, y = self.tile_index % s, self.tile_index // s
        vw, vh = self.width, self.height
        base, big = (vw // s, vh // s), (vw % s, vh % s)

        def index(gx, gy):
            return tuple(min(g, b) * (bs + 1) + max(g - b, 0) * bs for g, b, bs in zip((gx, gy), big, base))

        return index(x, y) + index(x + 1, y + 1)

    def render(self, display: DeviceBuffer | None = None, work_counters: DeviceBuffer | None = None):
        """Render (VolumetricCloud.cpp:537-566): one tile of one frame, then the display pass."""
        if self.tile_index == 0:
            self.frame_cnt += 1
            self.rendered_mask.zero()
        path_trace(self.ctx, self.atmosphere_block, self.render_block, self.common, self.material, self.params,
                   self.inputs, self.get_render_region(), self.frame_cnt, 1, self.accumulation, self.rendered_mask,
                   self.width, self.height, work_counters)
        if display is not None:
            self.display(display)
        self.tile_index = (self.tile_index + 1) % (self.sqrt_tile_count * self.sqrt_tile_count)

    Render = render

    def render_frames(self, frames: int, work_counters: DeviceBuffer | None = None):
        """`frames` whole frames (sqrt_tile_count 1) in one launch: the same samples and the
        same accumulation order as that many Render calls."""
        if self.sqrt_tile_count != 1 or self.tile_index != 0:
            raise ValueError("render_frames needs a single-tile schedule at a frame boundary")
        self.rendered_mask.zero()
        path_trace(self.ctx, self.atmosphere_block, self.render_block, self.common, self.material, self.params,
                   self.inputs, (0, 0, self.width, self.height), self.frame_cnt + 1, frames, self.accumulation,
                   self.rendered_mask, self.width, self.height, work_counters)
        self.frame_cnt += frames

    def display(self, display: DeviceBuffer):
        _abi.check(_abi.lib().skyhip_cloud_path_trace_display(self.ctx.handle, self.accumulation.ptr,
                                                              self.rendered_mask.ptr, self.frame_cnt, display.ptr,
                                                              self.width, self.height))

    def accumulated(self):
        return self.accumulation.download(np.float32, (self.height, self.width, 4))

    def mean(self):
        return self.accumulated() / max(self.frame_cnt, 1)


class PathTraceScene:
    """Everything the path-trace pass reads, on the device: atmosphere LUTs for the
    workload's camera and sun (T, MS, sky view, aerial perspective), the cloud blocks of
    VolumetricCloud::Update, the voxel grid with its majorants, and a PathTracing object."""

    def __init__(self, ctx: Context, width=None, height=None, workload=None, grid: VoxelGrid | None = None,
                 init: PathTracingInitParam | None = None, bricks=True):
        w = workload or presets.path_trace_workload()
        self.ctx = ctx
        self.workload = w
        self.width = width or w["width"]
        self.height = height or w["height"]
        cam: Camera = w["camera"]
        cam.aspect = self.width / self.height
        self.camera = cam
        self.earth = w["atmosphere"]
        self.atmosphere = Atmosphere(ctx)
        self.atmosphere.update_luts(self.earth)
        self.render_params = AtmosphereRenderParameters(
            sun_direction_theta=w["sun_theta"], sun_direction_phi=w["sun_phi"], sky_view_lut_steps=30.0,
            aerial_perspective_lut_steps=50.867000579833984, aerial_perspective_lut_max_distance=203.0,
            camera_position=cam.position, view_projection=cam.view_projection())
        self.renderer = AtmosphereRenderer(ctx, AtmosphereRenderInitParameters(sky_view_lut_size=(192, 108)))
        self.renderer.render(self.atmosphere, self.earth, self.render_params)
        self.cloud = w["cloud"]
        self.common, self.cloud_block, self.model = self.cloud.update(
            cam, self.earth.bottom_radius, self.renderer.sun_direction, 203.0, 0)
        # Grid placed ahead of the camera (local frame), footprint w["grid_extent_km"].
        if grid is None:
            grid = VoxelGrid(ctx, w["grid_dims"], brick=w.get("brick")).generate(w["grid"])
        self.grid = grid
        center = presets.grid_center_ahead(self.model, cam, w["grid_distance_km"])
        ext = w["grid_extent_km"]
        freq = (1.0 / ext, 1.0 / ext)
        bias = (0.5 - center[0] / ext, 0.5 - center[1] / ext)
        self.material = voxel_f16_material(grid, freq, bias, w["density"])
        self.sigma_t_max = grid.max_density() * w["density"]
        a, r = self.atmosphere, self.renderer
        i = _abi.PathTracingInputsC()
        i.transmittance = a.transmittance_texture.ptr
        i.transmittance_width, i.transmittance_height = a.transmittance_size
        i.sky_view_luminance = r.sky_view_luminance_texture.ptr
        i.sky_view_transmittance = r.sky_view_transmittance_texture.ptr
        i.sky_view_width, i.sky_view_height = r.init.sky_view_lut_size
        i.ap_luminance = r.aerial_perspective_luminance_texture.ptr
        i.ap_transmittance = r.aerial_perspective_transmittance_texture.ptr
        i.ap_width, i.ap_height = r.init.aerial_perspective_lut_size
        i.ap_depth = r.init.aerial_perspective_lut_depth
        i.brick_majorants = grid.majorants.ptr if bricks else None
        i.brick_size = grid.brick if bricks else 0
        i.cell_corners = grid.cells.ptr if (bricks and grid.cells is not None) else None
        self.inputs = i
        self.init = init or w["init"]
        self.tracer = PathTracing(ctx, (self.width, self.height), self.init, a.buffer, r.buffer, self.common,
                                  self.material, i, self.sigma_t_max, self.model)
        self.counter = ctx.alloc(5 * 8)

    def count_work(self, frames=1) -> dict:
        """Runs `frames` frames with the counting kernel variant on a scratch accumulator."""
        self.counter.zero()
        acc = self.ctx.alloc(self.width * self.height * 16)
        acc.zero()
        path_trace(self.ctx, self.atmosphere.buffer, self.renderer.buffer, self.common, self.material,
                   self.tracer.params, self.inputs, (0, 0, self.width, self.height), 1, frames, acc, None,
                   self.width, self.height, self.counter)
        c = self.counter.download(np.uint64, (5,))
        acc.free()
        return {"samples": int(c[0]), "lookups": int(c[1]), "segments": int(c[2]), "vertices": int(c[3]),
                "shadow_rays": int(c[4])}

    def host_inputs(self):
        """Host copies for the oracle: (T, sv_l, sv_t, ap_l, ap_t, grid bits)."""
        a, r = self.atmosphere, self.renderer
        sv_l, sv_t = r.sky_view()
        ap_l, ap_t = r.aerial_perspective()
        return a.transmittance(), sv_l, sv_t, ap_l, ap_t, self.grid.download()
