"""Scene assembly for the cloud raymarch: the atmosphere LUTs it reads, the material
textures, the std140 blocks and the output images, all resident on the device.

Mirrors the order of the reference's frame (AppWindow::HandleDisplayEvent ->
Earth::Update -> Atmosphere::UpdateLuts; AtmosphereRenderer::Render (sky view, aerial
perspective); VolumetricCloud::Update; VolumetricCloud::Render's raymarch pass).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi, presets
from .atmosphere import (Atmosphere, AtmosphereRenderer, AtmosphereRenderInitParameters,
                         AtmosphereRenderParameters)
from .cloud import CloudTextures, VolumetricCloudMinimalMaterial, bind_cloud_phase, cloud_raymarch, make_material
from .device import Context


class CloudScene:
    def __init__(self, ctx: Context, width=1920, height=1080, workload=None, material=None,
                 weather="uniform", blue_noise=None, frame_id=0, lod_bias=None, texture_cells=True):
        w = workload or presets.cloud_workload()
        self.ctx = ctx
        self.width, self.height = width, height
        self.cloud = w["cloud"]
        self.material = material or w["material"]
        if lod_bias is not None and hasattr(self.material, "lod_bias"):
            self.material.lod_bias = lod_bias
        cam = w["camera"]
        cam.aspect = width / height
        self.camera = cam
        self.earth = w["atmosphere"]
        # Atmosphere LUTs (T, MS) and the renderer's aerial-perspective LUT for this camera.
        self.atmosphere = Atmosphere(ctx)
        self.atmosphere.update_luts(self.earth)
        self.render_params = AtmosphereRenderParameters(
            sun_direction_theta=w["sun_theta"], sun_direction_phi=w["sun_phi"], sky_view_lut_steps=30.0,
            aerial_perspective_lut_steps=50.867000579833984, aerial_perspective_lut_max_distance=203.0,
            camera_position=cam.position, view_projection=cam.view_projection())
        self.renderer = AtmosphereRenderer(ctx, AtmosphereRenderInitParameters(sky_view_lut_size=(192, 108)))
        self.renderer.render(self.atmosphere, self.earth, self.render_params)
        # Material textures.
        self.textures = None
        if not isinstance(self.material, VolumetricCloudMinimalMaterial):
            m = self.material
            self.textures = CloudTextures(ctx, m.cloud_map_size, m.detail_size, m.displacement_size,
                                          worley_size=getattr(m, "worley_size", None), cells=texture_cells)
            self.textures.generate(m)
            if weather == "uniform":
                self.textures.upload_level0(cloud_map=presets.uniform_weather_map(m.cloud_map_size,
                                                                                  w["weather_seeds"]))
        # Blue noise.
        if blue_noise is None:
            if w.get("blue_noise_kind") == "blue":
                blue_noise = presets.blue_noise_tile(w["blue_noise_size"], w["blue_noise_seed"])
            else:
                blue_noise = presets.uniform_blue_noise(w["blue_noise_size"], w["blue_noise_seed"])
        # The raymarch's phase lobes (None: the reference's), bound on the context for each run.
        self.phase = w.get("phase")
        self.blue_noise_host = np.ascontiguousarray(blue_noise, dtype=np.uint16)
        self.blue_noise = ctx.upload(self.blue_noise_host)
        # Blocks (VolumetricCloud::Update).
        self.common, self.cloud_block, self.model = self.cloud.update(
            cam, self.earth.bottom_radius, self.renderer.sun_direction, 203.0, frame_id)
        self.mat_block = make_material(self.material, self.textures, (width, height), cam)
        # Outputs.
        self.out_rgba = ctx.alloc(width * height * 8)
        self.out_dist = ctx.alloc(width * height * 4)
        self.counter = ctx.alloc(32)
        self.pixel_steps = ctx.alloc(width * height * 8)

    def run(self, count_steps=False, depth=None, shading_index=None):
        if count_steps:
            self.counter.zero()
        a = self.atmosphere
        r = self.renderer
        aw, ah = r.init.aerial_perspective_lut_size
        bind_cloud_phase(self.ctx, self.phase)
        cloud_raymarch(self.ctx, a.buffer, self.common, self.cloud_block, self.mat_block, a.transmittance_texture,
                       a.transmittance_size, r.aerial_perspective_luminance_texture,
                       r.aerial_perspective_transmittance_texture, (aw, ah, r.init.aerial_perspective_lut_depth),
                       self.blue_noise, self.blue_noise_host.shape[0], self.out_rgba, self.out_dist, self.width,
                       self.height, depth=depth, shading_index=shading_index,
                       work_counters=self.counter if count_steps else None,
                       pixel_steps=self.pixel_steps if count_steps else None)

    def executed_steps(self) -> int:
        return int(self.counter.download(np.uint64, (4,))[0])

    def count_work(self, **kw) -> dict:
        """Runs the counting variant once: executed primary steps, density evaluations,
        detail fetches and in-cloud steps of one frame."""
        self.run(count_steps=True, **kw)
        c = self.counter.download(np.uint64, (4,))
        return {"steps": int(c[0]), "evals": int(c[1]), "detail": int(c[2]), "cloud_steps": int(c[3])}

    def steps_per_pixel(self):
        """Per-pixel step records of the last counting run, uint32[n, 2]: executed and
        scheduled steps, each packed segment 0 | segment 1 << 16."""
        return self.pixel_steps.download(np.uint32, (self.height * self.width, 2))

    def image(self):
        rgba = self.out_rgba.download(np.float16, (self.height, self.width, 4)).astype(np.float32)
        dist = self.out_dist.download(np.float32, (self.height, self.width))
        return rgba, dist
