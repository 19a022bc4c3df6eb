"""Host mirror of the reference's volumetric-cloud classes over the skyhip C ABI.

Names, fields and defaults follow src/SkyRendering/VolumetricCloud.h:16-56 and
VolumetricCloudDefaultMaterial.h:57-207 (the material plugin of
IVolumetricCloudMaterial.h:9-30 becomes a tagged parameter set: SkyhipCloudMaterial).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields

import numpy as np

from . import _abi
from .atmosphere import Camera
from .device import Context, DeviceBuffer

u8p = C.POINTER(C.c_uint8)
_abi.register("skyhip_mip_levels", C.c_int, [C.c_int32, C.c_int32, C.c_int32])
_abi.register("skyhip_mip_chain_bytes", C.c_size_t, [C.c_int32] * 5)
_abi.register("skyhip_noise_cloud_map", C.c_int, [C.c_void_p, C.POINTER(_abi.NoiseCreateInfo),
                                                   C.POINTER(_abi.NoiseCreateInfo), C.c_void_p, C.c_int32])
_abi.register("skyhip_noise_detail", C.c_int, [C.c_void_p, C.POINTER(_abi.NoiseCreateInfo),
                                                C.POINTER(_abi.NoiseCreateInfo), C.c_void_p, C.c_int32])
_abi.register("skyhip_noise_displacement", C.c_int, [C.c_void_p, C.POINTER(_abi.NoiseCreateInfo), C.c_void_p,
                                                      C.c_int32])
_abi.register("skyhip_noise_worley", C.c_int, [C.c_void_p, C.POINTER(_abi.NoiseCreateInfo), C.c_void_p, C.c_int32])
_abi.register("skyhip_generate_mipmaps", C.c_int, [C.c_void_p, C.c_void_p] + [C.c_int32] * 6)
_abi.register("skyhip_texture_cells_2d_rg8", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p])
_abi.register("skyhip_texture_cells_3d_r8", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p])
_abi.register("skyhip_cloud_raymarch", C.c_int, [
    C.c_void_p, C.POINTER(_abi.AtmosphereBufferData), C.POINTER(_abi.CloudCommonBufferData),
    C.POINTER(_abi.CloudBufferData), C.POINTER(_abi.CloudMaterial), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
    C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
    C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p])


_abi.register("skyhip_cloud_sample_sigma_t", C.c_int, [C.c_void_p, C.POINTER(_abi.CloudCommonBufferData),
                                                       C.POINTER(_abi.CloudMaterial), C.c_void_p, C.c_int32,
                                                       C.c_void_p])


def sample_sigma_t(ctx: Context, common, material: _abi.CloudMaterial, positions: np.ndarray) -> np.ndarray:
    """The material's SampleSigmaT at local-frame points (n x 3), evaluated on the device."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    dpos = ctx.upload(pos)
    out = ctx.alloc(max(4 * len(pos), 4))
    _abi.check(_abi.lib().skyhip_cloud_sample_sigma_t(ctx.handle, C.byref(common), C.byref(material), dpos.ptr,
                                                      len(pos), out.ptr))
    return out.download(np.float32, (len(pos),))


def mip_levels(w, h=1, d=1) -> int:
    """GetMipmapLevels, src/Base/include/ImageLoader.h:8-14."""
    return int(_abi.lib().skyhip_mip_levels(w, h, d))


def mip_chain_bytes(w, h, d, channels, levels) -> int:
    return int(_abi.lib().skyhip_mip_chain_bytes(w, h, d, channels, levels))


@dataclass
class NoiseCreateInfo:
    """NoiseCreateInfo, VolumetricCloudDefaultMaterial.h:57-62."""
    seed: int = 0
    base_frequency: int = 1
    remap_min: float = 0.0
    remap_max: float = 1.0

    def to_c(self):
        return _abi.NoiseCreateInfo(self.seed & 0xFFFFFFFF, self.base_frequency, self.remap_min, self.remap_max)


@dataclass
class DefaultMaterialCommon:
    """VolumetricCloudDefaultMaterialCommon, VolumetricCloudDefaultMaterial.h:86-140 and .cpp:30-76."""
    cloud_map_density: NoiseCreateInfo = field(default_factory=lambda: NoiseCreateInfo(0, 3, 0.35, 0.75))
    cloud_map_height: NoiseCreateInfo = field(default_factory=lambda: NoiseCreateInfo(0, 5, 0.8, 0.4))
    detail_perlin: NoiseCreateInfo = field(default_factory=lambda: NoiseCreateInfo(0, 7, 0.23, 1.0))
    detail_worley: NoiseCreateInfo = field(default_factory=lambda: NoiseCreateInfo(0, 11, 1.0, 0.0))
    displacement_perlin: NoiseCreateInfo = field(default_factory=lambda: NoiseCreateInfo(0, 6, 0.25, 0.75))
    cloud_map_repeat_size: float = 18.99
    detail_repeat_size: float = 5.33
    displacement_repeat_size: float = 3.51
    lod_bias: float = 2.75
    density: float = 15.0
    wind_speed: float = 0.05
    detail_wind_magnify: float = 1.0
    cloud_map_size: int = 512
    detail_size: int = 128
    displacement_size: int = 128

    def load_json(self, d: dict):
        """The `materail_common_` object of bin/config*.json (flattened FIELD_DECLARE keys)."""
        noise = {"cloud_map_.buffer.uDensity": "cloud_map_density", "cloud_map_.buffer.uHeight": "cloud_map_height",
                 "detail_.buffer.uPerlin": "detail_perlin", "detail_.buffer.uWorley": "detail_worley",
                 "displacement_.buffer.uPerlin": "displacement_perlin"}
        for prefix, attr in noise.items():
            n = getattr(self, attr)
            for k in ("seed", "base_frequency", "remap_min", "remap_max"):
                if f"{prefix}.{k}" in d:
                    setattr(n, k, type(getattr(n, k))(d[f"{prefix}.{k}"]))
        plain = {"cloud_map_.texture.repeat_size": "cloud_map_repeat_size",
                 "detail_.texture.repeat_size": "detail_repeat_size",
                 "displacement_.texture.repeat_size": "displacement_repeat_size", "density_": "density",
                 "lod_bias_": "lod_bias", "wind_speed_": "wind_speed", "detail_wind_magnify_": "detail_wind_magnify"}
        for k, attr in plain.items():
            if k in d:
                setattr(self, attr, float(d[k]))
        return self

    def common_block(self, viewport, camera: Camera, offset=(0.0, 0.0), detail_offset=(0.0, 0.0)):
        """VolumetricCloudDefaultMaterialCommon::Update (VolumetricCloudDefaultMaterial.cpp:78-102) for a
        static frame (additional wind delta 0)."""
        blk = _abi.CloudMaterialCommonBufferData()
        blk.uLodBias = self.lod_bias
        blk.uDensity = self.density
        L = _abi.lib()
        vw, vh = viewport
        for info, repeat, width, off in [
                (blk.uCloudMapSampleInfo, self.cloud_map_repeat_size, self.cloud_map_size, offset),
                (blk.uDetailSampleInfo, self.detail_repeat_size, self.detail_size,
                 (offset[0] + detail_offset[0], offset[1] + detail_offset[1])),
                (blk.uDisplacementSampleInfo, self.displacement_repeat_size, self.displacement_size, offset)]:
            _abi.check(L.skyhip_cloud_sample_info(repeat, width, camera.fovy, vw, vh, off[0], off[1], C.byref(info)))
        return blk


@dataclass
class VolumetricCloudDefaultMaterial1(DefaultMaterialCommon):
    """VolumetricCloudDefaultMaterial1, VolumetricCloudDefaultMaterial.h:171-207, .cpp:224-257."""
    detail_base: float = 0.67
    detail_scale: float = 1.86
    base_density_threshold: float = 0.4
    base_height_hardness: float = 6.0
    base_edge_hardness: float = 6.0
    height_cut: float = 0.9
    edge_cut: float = 0.8
    kind = _abi.MATERIAL_DEFAULT1

    @classmethod
    def from_json(cls, data: dict):
        m = cls()
        m.load_json(data.get("materail_common_", {}))
        for k in ("detail_base", "detail_scale", "base_density_threshold", "base_height_hardness",
                  "base_edge_hardness", "height_cut", "edge_cut"):
            if k + "_" in data:
                setattr(m, k, float(data[k + "_"]))
        return m

    def material1_block(self):
        b = _abi.CloudMaterial1BufferData()
        b.uBaseDensityThreshold = self.base_density_threshold
        b.uBaseHeightHardness = self.base_height_hardness
        b.uBaseEdgeHardness = self.base_edge_hardness
        b.uDetailBase = self.detail_base
        b.uDetailScale = self.detail_scale
        b.uHeightCut = np.float32(1.0) - np.float32(self.height_cut)  # .cpp:253
        b.uEdgeCur = self.edge_cut
        return b


@dataclass
class VolumetricCloudDefaultMaterial1Worley(VolumetricCloudDefaultMaterial1):
    """Material1 with a second, high-frequency Worley detail volume (SKYHIP_MATERIAL_DEFAULT1_WORLEY;
    no reference counterpart).  The volume is WorleyFBM alone (worley_size^3 R8, mips), tiled
    every worley_repeat_size km; it erodes Material1's density by
    (worley + worley_base) * worley_scale * Material1's erosion mask."""
    worley: NoiseCreateInfo = field(default_factory=lambda: NoiseCreateInfo(4, 2, 0.0, 1.0))
    worley_size: int = 32
    worley_repeat_size: float = 1.5
    worley_base: float = 0.0
    worley_scale: float = 0.35
    kind = _abi.MATERIAL_DEFAULT1_WORLEY

    @classmethod
    def from_material1(cls, m1: VolumetricCloudDefaultMaterial1, **kw):
        m = cls(**{f.name: getattr(m1, f.name) for f in fields(m1)})
        for k, v in kw.items():
            setattr(m, k, v)
        return m

    def worley_block(self, viewport, camera: Camera, offset=(0.0, 0.0)):
        """The Worley block; its SampleInfo is formed as the default material's (CalKLod,
        VolumetricCloudDefaultMaterial.cpp:24-28, 78-102) for the Worley volume's size and tiling."""
        b = _abi.CloudWorleyDetailBufferData()
        vw, vh = viewport
        _abi.check(_abi.lib().skyhip_cloud_sample_info(self.worley_repeat_size, self.worley_size, camera.fovy, vw, vh,
                                                       offset[0], offset[1], C.byref(b.uWorleySampleInfo)))
        b.uWorleyBase = self.worley_base
        b.uWorleyScale = self.worley_scale
        return b


@dataclass
class VolumetricCloudDefaultMaterial0(DefaultMaterialCommon):
    """VolumetricCloudDefaultMaterial0, VolumetricCloudDefaultMaterial.h:142-169, .cpp:180-222."""
    detail_param: tuple = (0.4, 0.0)
    displacement_scale: float = 1.0
    kind = _abi.MATERIAL_DEFAULT0

    @classmethod
    def from_json(cls, data: dict):
        m = cls()
        m.load_json(data.get("materail_common_", {}))
        if "detail_param_" in data:
            m.detail_param = tuple(float(x) for x in data["detail_param_"])
        if "displacement_scale_" in data:
            m.displacement_scale = float(data["displacement_scale_"])
        return m

    def material0_block(self):
        b = _abi.CloudMaterial0BufferData()  # Material0::Update (.cpp:198-205)
        b.uDetailParam[0], b.uDetailParam[1] = self.detail_param
        b.uDisplacementScale = self.displacement_scale
        return b


@dataclass
class VolumetricCloudVoxelMaterial:
    """VolumetricCloudVoxelMaterial, VolumetricCloudVoxelMaterial.h:10-40 (an R8 voxel volume
    with mips, placed by base_ and width_ in km)."""
    lod_bias: float = 2.75
    density: float = 20.0
    base: tuple = (0.0, 0.0)
    width: tuple = (2.0, 2.0)
    kind = _abi.MATERIAL_VOXEL

    def voxel_block(self, viewport, camera: Camera, voxel_dim, offset=(0.0, 0.0)):
        """VolumetricCloudVoxelMaterial::Update (VolumetricCloudVoxelMaterial.cpp:93-104)."""
        b = _abi.CloudVoxelBufferData()
        f32 = np.float32
        b.uLodBias, b.uDensity = self.lod_bias, self.density
        for k in range(2):
            b.uSampleFrequency[k] = f32(1.0) / f32(self.width[k])
            b.uSampleBias[k] = f32((np.float64(offset[k]) + np.float64(f32(self.base[k]))) / np.float64(f32(self.width[k])))
        max_width = f32(max(voxel_dim))
        tan_half = f32(np.tan(f32(np.radians(f32(camera.fovy))) * f32(0.5)))
        b.uSampleLodK = max_width * tan_half / (f32(min(self.width)) * f32(min(viewport)))
        return b


@dataclass
class VolumetricCloudMinimalMaterial:
    """VolumetricCloudMinimalMaterial (constant extinction), VolumetricCloudMinimalMaterial.h."""
    density: float = 0.5
    kind = _abi.MATERIAL_MINIMAL


class CloudTextures:
    """Device mip chains of the default material's textures (cloud map RG8 2D, detail R8 3D,
    displacement RGBA8 2D, optional Worley R8 3D), generated by the noise kernels or uploaded
    from the host, and with `cells` the cell-corner copies of the cloud map's, detail's and
    Worley volume's level 0 (one 8-byte load per LINEAR fetch; 4x / 8x / 8x level 0)."""

    def __init__(self, ctx: Context, cloud_map_size=512, detail_size=128, displacement_size=128, worley_size=None,
                 cells=True):
        self.ctx = ctx
        self.cells = cells
        self.cloud_map_size, self.detail_size, self.displacement_size = cloud_map_size, detail_size, displacement_size
        self.worley_size = worley_size
        self.worley = None
        if worley_size:
            self.worley_levels = mip_levels(worley_size, worley_size, worley_size)
            self.worley = ctx.alloc(mip_chain_bytes(worley_size, worley_size, worley_size, 1, self.worley_levels))
        self.cloud_map_levels = mip_levels(cloud_map_size, cloud_map_size)
        self.detail_levels = mip_levels(detail_size, detail_size, detail_size)
        self.displacement_levels = mip_levels(displacement_size, displacement_size)
        self.cloud_map = ctx.alloc(mip_chain_bytes(cloud_map_size, cloud_map_size, 1, 2, self.cloud_map_levels))
        self.detail = ctx.alloc(mip_chain_bytes(detail_size, detail_size, detail_size, 1, self.detail_levels))
        self.displacement = ctx.alloc(
            mip_chain_bytes(displacement_size, displacement_size, 1, 4, self.displacement_levels))
        self.cloud_map_cells = ctx.alloc(cloud_map_size ** 2 * 8) if cells else None
        self.detail_cells = ctx.alloc(detail_size ** 3 * 8) if cells else None
        self.worley_cells = ctx.alloc(worley_size ** 3 * 8) if (cells and worley_size) else None

    def generate(self, m: DefaultMaterialCommon):
        """DynamicTexture::Generate for all three textures (VolumetricCloudDefaultMaterial.h:39-49)."""
        L = _abi.lib()
        h = self.ctx.handle
        _abi.check(L.skyhip_noise_cloud_map(h, C.byref(m.cloud_map_density.to_c()), C.byref(m.cloud_map_height.to_c()),
                                            self.cloud_map.ptr, self.cloud_map_size))
        _abi.check(L.skyhip_noise_detail(h, C.byref(m.detail_perlin.to_c()), C.byref(m.detail_worley.to_c()),
                                         self.detail.ptr, self.detail_size))
        _abi.check(L.skyhip_noise_displacement(h, C.byref(m.displacement_perlin.to_c()), self.displacement.ptr,
                                               self.displacement_size))
        if self.worley is not None:
            _abi.check(L.skyhip_noise_worley(h, C.byref(m.worley.to_c()), self.worley.ptr, self.worley_size))
        self.build_mips()
        return self

    def upload_level0(self, cloud_map=None, detail=None, displacement=None):
        if cloud_map is not None:
            self.cloud_map.upload(np.ascontiguousarray(cloud_map, dtype=np.uint8))
        if detail is not None:
            self.detail.upload(np.ascontiguousarray(detail, dtype=np.uint8))
        if displacement is not None:
            self.displacement.upload(np.ascontiguousarray(displacement, dtype=np.uint8))
        self.build_mips()
        return self

    def build_mips(self):
        L = _abi.lib()
        h = self.ctx.handle
        n = self.cloud_map_size
        _abi.check(L.skyhip_generate_mipmaps(h, self.cloud_map.ptr, 2, 2, n, n, 1, self.cloud_map_levels))
        n = self.detail_size
        _abi.check(L.skyhip_generate_mipmaps(h, self.detail.ptr, 3, 1, n, n, n, self.detail_levels))
        n = self.displacement_size
        _abi.check(L.skyhip_generate_mipmaps(h, self.displacement.ptr, 2, 4, n, n, 1, self.displacement_levels))
        if self.worley is not None:
            n = self.worley_size
            _abi.check(L.skyhip_generate_mipmaps(h, self.worley.ptr, 3, 1, n, n, n, self.worley_levels))
        self.build_cells()

    def build_cells(self):
        """The level-0 cell-corner copies (kept in step with the textures by build_mips)."""
        if not self.cells:
            return
        L = _abi.lib()
        h = self.ctx.handle
        _abi.check(L.skyhip_texture_cells_2d_rg8(h, self.cloud_map.ptr, self.cloud_map_size, self.cloud_map_cells.ptr))
        _abi.check(L.skyhip_texture_cells_3d_r8(h, self.detail.ptr, self.detail_size, self.detail_cells.ptr))
        if self.worley_cells is not None:
            _abi.check(L.skyhip_texture_cells_3d_r8(h, self.worley.ptr, self.worley_size, self.worley_cells.ptr))

    def download(self):
        return (self.cloud_map.download(np.uint8, (self.cloud_map.nbytes,)),
                self.detail.download(np.uint8, (self.detail.nbytes,)),
                self.displacement.download(np.uint8, (self.displacement.nbytes,)))

    def download_worley(self):
        return None if self.worley is None else self.worley.download(np.uint8, (self.worley.nbytes,))

    def fill_material(self, mat: _abi.CloudMaterial):
        mat.cloud_map, mat.cloud_map_size, mat.cloud_map_levels = self.cloud_map.ptr, self.cloud_map_size, \
            self.cloud_map_levels
        mat.detail, mat.detail_size, mat.detail_levels = self.detail.ptr, self.detail_size, self.detail_levels
        mat.displacement, mat.displacement_size, mat.displacement_levels = self.displacement.ptr, \
            self.displacement_size, self.displacement_levels
        if self.worley is not None:
            mat.worley_texture, mat.worley_size, mat.worley_levels = self.worley.ptr, self.worley_size, \
                self.worley_levels
        if self.cells:
            mat.cloud_map_cells, mat.detail_cells = self.cloud_map_cells.ptr, self.detail_cells.ptr
            mat.worley_cells = self.worley_cells.ptr if self.worley_cells is not None else None


@dataclass
class VolumetricCloud:
    """VolumetricCloud's serialised fields (VolumetricCloud.h:20-35) and its Update/Render
    (VolumetricCloud.cpp:168-280, 327-386) for the raymarch pass."""
    bottom_altitude: float = 2.0
    thickness: float = 2.0
    max_raymarch_distance: float = 30.0
    max_raymarch_steps: float = 128.0
    max_visible_distance: float = 120.0
    env_color: tuple = (1.0, 1.0, 1.0)
    env_color_scale: float = 0.1
    sun_illuminance_scale: float = 1.0
    shadow_steps: float = 5.0
    shadow_distance: float = 2.0
    shadow_map_max_distance: float = 20.0
    shadow_froxel_max_distance: float = 20.0
    sun_multiscattering_sigma_scale: float = 0.3
    env_multiscattering_sigma_scale: float = 0.5
    env_bottom_visibility: float = 0.4
    env_sun_height_curve_exp: float = 1.0

    @classmethod
    def from_json(cls, d: dict):
        v = cls()
        for f in fields(cls):
            k = f.name + "_" if (f.name + "_") in d else f.name
            if k in d:
                setattr(v, f.name, tuple(d[k]) if isinstance(d[k], list) else float(d[k]))
        return v

    def to_c(self):
        c = _abi.VolumetricCloudParametersC()
        for f in fields(self):
            val = getattr(self, f.name)
            if isinstance(val, tuple):
                for i, x in enumerate(val):
                    c.env_color[i] = x
            else:
                setattr(c, f.name, val)
        return c

    def update(self, camera: Camera, bottom_radius: float, sun_direction_world, ap_max_distance: float,
               frame_id: int = 0):
        """Returns (common block, cloud block, model matrix) as VolumetricCloud::Update fills them."""
        common, cloud = _abi.CloudCommonBufferData(), _abi.CloudBufferData()
        model = (C.c_float * 16)()
        sun = (C.c_float * 3)(*[float(x) for x in sun_direction_world])
        _abi.check(_abi.lib().skyhip_cloud_assign_buffer_data(C.byref(self.to_c()), C.byref(camera.to_c()),
                                                              float(bottom_radius), sun, float(ap_max_distance),
                                                              int(frame_id), C.byref(common), C.byref(cloud), model))
        return common, cloud, np.array(model[:], dtype=np.float32)


_abi.register("skyhip_cloud_shadow_matrices", C.c_int, [
    C.POINTER(_abi.VolumetricCloudParametersC), C.POINTER(_abi.CameraC), C.c_float, C.c_void_p, C.c_int32, C.c_int32,
    C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(_abi.CloudCommonBufferData), C.c_void_p, C.c_void_p])
_abi.register("skyhip_cloud_shadow_map", C.c_int, [
    C.c_void_p, C.POINTER(_abi.CloudCommonBufferData), C.POINTER(_abi.CloudMaterial), C.c_void_p, C.c_int32,
    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32])
_abi.register("skyhip_cloud_shadow_froxel", C.c_int, [
    C.c_void_p, C.POINTER(_abi.CloudCommonBufferData), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
    C.c_int32, C.c_int32])

_abi.register("skyhip_bind_shadow_froxel", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32])


def bind_shadow_froxel(ctx: Context, froxel: DeviceBuffer | None, size=(0, 0, 0)):
    """skyhip_bind_shadow_froxel: froxel (R16 UNORM, size = (w, h, d)) multiplies the
    aerial-perspective term of later raymarch and path-trace calls; None unbinds."""
    w, h, d = size
    _abi.check(_abi.lib().skyhip_bind_shadow_froxel(ctx.handle, froxel.ptr if froxel is not None else None, w, h, d))


_abi.register("skyhip_bind_cloud_phase", C.c_int, [C.c_void_p, C.POINTER(_abi.CloudPhase)])

REFERENCE_PHASE = (-0.15, 2.16, 0.85)  # VolumetricCloudRender.comp:125-126: (back g, back scale, forward g)


def bind_cloud_phase(ctx: Context, phase=None):
    """skyhip_bind_cloud_phase: the lobes (back_g, back_scale, forward_g) of later raymarch calls
    on ctx, phase = mix(HG(c, back_g) * back_scale, HG(c, forward_g), exp(-s sigma)); None
    restores the reference's lobes."""
    if phase is None:
        _abi.check(_abi.lib().skyhip_bind_cloud_phase(ctx.handle, None))
    else:
        blk = _abi.CloudPhase(*[float(x) for x in phase], 0.0)
        _abi.check(_abi.lib().skyhip_bind_cloud_phase(ctx.handle, C.byref(blk)))


SHADOW_MAP_RESOLUTION = (512, 512)  # kShadowMapResolution, VolumetricCloud.cpp:52
SHADOW_FROXEL_DEPTH = 128           # VolumetricCloud.cpp:135


def cloud_shadow_matrices(cloud: "VolumetricCloud", camera: Camera, bottom_radius: float, sun_direction_world,
                          common, previous_light_vp=None, previous_model=None, additional_delta=None,
                          shadow_map_size=SHADOW_MAP_RESOLUTION):
    """Fills common's light matrices as VolumetricCloud::Update does; returns (light VP,
    light VP * model^-1) as float32 [16] column-major arrays."""
    def arr(x, n):
        return None if x is None else (C.c_float * n)(*[float(v) for v in np.asarray(x, np.float32).ravel()])
    out_vp, out_vpm = (C.c_float * 16)(), (C.c_float * 16)()
    _abi.check(_abi.lib().skyhip_cloud_shadow_matrices(
        C.byref(cloud.to_c()), C.byref(camera.to_c()), float(bottom_radius), arr(sun_direction_world, 3),
        shadow_map_size[0], shadow_map_size[1], arr(previous_light_vp, 16), arr(previous_model, 16),
        arr(additional_delta, 2), C.byref(common), out_vp, out_vpm))
    return np.array(out_vp[:], np.float32), np.array(out_vpm[:], np.float32)


class CloudShadow:
    """VolumetricCloud::RenderShadow (VolumetricCloud.cpp:282-325): the raw shadow maps of this
    and the previous frame, the blurred map and the shadow froxel for a viewport."""

    def __init__(self, ctx: Context, viewport, map_size=SHADOW_MAP_RESOLUTION, froxel_depth=SHADOW_FROXEL_DEPTH):
        self.ctx = ctx
        self.map_size = tuple(map_size)
        mw, mh = self.map_size
        self.raw = [ctx.alloc(mw * mh * 8), ctx.alloc(mw * mh * 8)]  # this frame's, the previous frame's
        self.scratch = ctx.alloc(mw * mh * 8)
        self.blurred = ctx.alloc(mw * mh * 8)
        self.froxel_size = (viewport[0] // 12, viewport[1] // 12, froxel_depth)
        fw, fh, fd = self.froxel_size
        self.froxel = ctx.alloc(fw * fh * fd * 2)
        self.frames = 0

    def render(self, common, material: _abi.CloudMaterial, blue_noise: DeviceBuffer, blue_noise_size: int):
        """One frame: the map (blended with the previous frame's after the first), blur, froxel."""
        self.raw.reverse()
        mw, mh = self.map_size
        previous = self.raw[1].ptr if self.frames > 0 else None
        _abi.check(_abi.lib().skyhip_cloud_shadow_map(
            self.ctx.handle, C.byref(common), C.byref(material), blue_noise.ptr, blue_noise_size, previous,
            self.raw[0].ptr, self.scratch.ptr, self.blurred.ptr, mw, mh))
        fw, fh, fd = self.froxel_size
        _abi.check(_abi.lib().skyhip_cloud_shadow_froxel(
            self.ctx.handle, C.byref(common), self.blurred.ptr, mw, mh, self.froxel.ptr, fw, fh, fd))
        self.frames += 1

    RenderShadow = render

    def bind(self):
        """Binds this froxel for the context's raymarch and path-trace calls (the reference binds
        GetShadowFroxel() to both passes, VolumetricCloud.cpp:369, 543)."""
        bind_shadow_froxel(self.ctx, self.froxel, self.froxel_size)

    def maps(self):
        mw, mh = self.map_size
        return (self.raw[0].download(np.float32, (mh, mw, 2)), self.blurred.download(np.float32, (mh, mw, 2)))

    def froxel_volume(self):
        fw, fh, fd = self.froxel_size
        return self.froxel.download(np.uint16, (fd, fh, fw))


def make_material(material, textures: CloudTextures | None, viewport, camera: Camera, offset=(0.0, 0.0)):
    m = _abi.CloudMaterial()
    m.kind = material.kind
    if material.kind == _abi.MATERIAL_MINIMAL:
        m.minimal.uDensity = material.density
        return m
    if material.kind == _abi.MATERIAL_VOXEL:
        raise ValueError("voxel materials need their volume: use voxel_material()")
    m.common = material.common_block(viewport, camera, offset)
    if material.kind == _abi.MATERIAL_DEFAULT0:
        m.material0 = material.material0_block()
    else:
        m.material1 = material.material1_block()
    if material.kind == _abi.MATERIAL_DEFAULT1_WORLEY:
        if textures.worley is None:
            raise ValueError("the Worley material needs CloudTextures with a Worley volume (worley_size)")
        m.worley = material.worley_block(viewport, camera, offset)
    textures.fill_material(m)
    return m


def voxel_material(material: "VolumetricCloudVoxelMaterial", voxel_texture: DeviceBuffer, voxel_dim, levels,
                   viewport, camera: Camera, offset=(0.0, 0.0)) -> _abi.CloudMaterial:
    """A VOXEL material over an R8 voxel mip chain (level 0 first, x fastest)."""
    m = _abi.CloudMaterial()
    m.kind = _abi.MATERIAL_VOXEL
    m.voxel = material.voxel_block(viewport, camera, voxel_dim, offset)
    m.voxel_texture = voxel_texture.ptr
    for k in range(3):
        m.voxel_dim[k] = int(voxel_dim[k])
    m.voxel_levels = int(levels)
    return m


def cloud_raymarch(ctx: Context, atmosphere_block, common, cloud, material: _abi.CloudMaterial,
                   transmittance: DeviceBuffer, t_size, ap_l: DeviceBuffer, ap_t: DeviceBuffer, ap_size,
                   blue_noise: DeviceBuffer, blue_noise_size, out_rgba16f: DeviceBuffer, out_distance: DeviceBuffer,
                   width, height, depth: DeviceBuffer | None = None, shading_index: DeviceBuffer | None = None,
                   work_counters: DeviceBuffer | None = None, pixel_steps: DeviceBuffer | None = None):
    """skyhip_cloud_raymarch: the VolumetricCloud render pass."""
    _abi.check(_abi.lib().skyhip_cloud_raymarch(
        ctx.handle, C.byref(atmosphere_block), C.byref(common), C.byref(cloud), C.byref(material), transmittance.ptr,
        t_size[0], t_size[1], ap_l.ptr, ap_t.ptr, ap_size[0], ap_size[1], ap_size[2], blue_noise.ptr, blue_noise_size,
        depth.ptr if depth is not None else None, shading_index.ptr if shading_index is not None else None,
        out_rgba16f.ptr, out_distance.ptr, width, height, work_counters.ptr if work_counters is not None else None,
        pixel_steps.ptr if pixel_steps is not None else None))


_abi.register("skyhip_cloud_reproject_matrix", C.c_int, [
    C.POINTER(_abi.CameraC), C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(_abi.CloudCommonBufferData),
    C.c_void_p])
_abi.register("skyhip_cloud_checkerboard_depth", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p])
_abi.register("skyhip_cloud_shading_index", C.c_int, [
    C.c_void_p, C.POINTER(_abi.CloudCommonBufferData), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p])
_abi.register("skyhip_cloud_reconstruct", C.c_int, [
    C.c_void_p, C.POINTER(_abi.CloudCommonBufferData), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
    C.c_void_p, C.c_int32, C.c_int32])
_abi.register("skyhip_cloud_upscale", C.c_int, [
    C.c_void_p, C.POINTER(_abi.CloudCommonBufferData), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
    C.c_int32])
_abi.register("skyhip_cloud_checkerboard_index", C.c_int, [
    C.c_void_p, C.POINTER(_abi.CloudCommonBufferData), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p])
_abi.register("skyhip_cloud_reconstruct_upscale", C.c_int, [
    C.c_void_p, C.POINTER(_abi.CloudCommonBufferData), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
    C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32])


def cloud_reproject_matrix(camera: Camera, bottom_radius: float, common, previous_mvp=None,
                           previous_camera_position=None, additional_delta=None):
    """Fills common.uReprojectMat as VolumetricCloud::Update does (VolumetricCloud.cpp:204-214,
    234); returns this frame's MVP (float32 [16], column-major), the next frame's previous_mvp."""
    def arr(x, n):
        return None if x is None else (C.c_float * n)(*[float(v) for v in np.asarray(x, np.float32).ravel()])
    out = (C.c_float * 16)()
    _abi.check(_abi.lib().skyhip_cloud_reproject_matrix(
        C.byref(camera.to_c()), float(bottom_radius), arr(previous_mvp, 16), arr(previous_camera_position, 3),
        arr(additional_delta, 2), C.byref(common), out))
    return np.array(out[:], np.float32)


class CloudReconstruction:
    """VolumetricCloud::Render's passes for one viewport (VolumetricCloud.cpp:327-422): the
    checkerboard depth (viewport / 2), the shading index (viewport / 4), the quarter-resolution
    raymarch of the indexed texels, the temporal reconstruction (viewport / 2, two textures
    that swap each frame) and the upscale composited into the HDR image."""

    def __init__(self, ctx: Context, viewport):
        w, h = int(viewport[0]), int(viewport[1])
        if w % 4 or h % 4:
            raise ValueError("the viewport's width and height must be multiples of 4")
        self.ctx, self.viewport = ctx, (w, h)
        self.checkerboard = ctx.alloc((w // 2) * (h // 2) * 4)
        self.index = ctx.alloc((w // 4) * (h // 4) * 8)
        self.render_texture = ctx.alloc((w // 4) * (h // 4) * 8)
        self.distance = ctx.alloc((w // 4) * (h // 4) * 4)
        self.reconstruct = [ctx.alloc((w // 2) * (h // 2) * 8), ctx.alloc((w // 2) * (h // 2) * 8)]
        self.frames = 0

    def checkerboard_depth(self, depth: DeviceBuffer):
        w, h = self.viewport
        _abi.check(_abi.lib().skyhip_cloud_checkerboard_depth(self.ctx.handle, depth.ptr, w, h, self.checkerboard.ptr))

    def shading_index(self, common):
        w, h = self.viewport
        _abi.check(_abi.lib().skyhip_cloud_shading_index(self.ctx.handle, C.byref(common), self.checkerboard.ptr, w,
                                                         h, self.index.ptr))

    def reconstruct_pass(self, common):
        w, h = self.viewport
        previous = self.reconstruct[1].ptr if self.frames > 0 else None
        _abi.check(_abi.lib().skyhip_cloud_reconstruct(
            self.ctx.handle, C.byref(common), self.checkerboard.ptr, self.index.ptr, self.render_texture.ptr,
            self.distance.ptr, previous, self.reconstruct[0].ptr, w, h))

    def upscale(self, common, depth: DeviceBuffer, hdr: DeviceBuffer):
        w, h = self.viewport
        _abi.check(_abi.lib().skyhip_cloud_upscale(self.ctx.handle, C.byref(common), self.checkerboard.ptr, depth.ptr,
                                                   self.reconstruct[0].ptr, hdr.ptr, w, h))

    def checkerboard_index(self, common, depth: DeviceBuffer):
        """checkerboard_depth + shading_index in one launch."""
        w, h = self.viewport
        _abi.check(_abi.lib().skyhip_cloud_checkerboard_index(self.ctx.handle, C.byref(common), depth.ptr, w, h,
                                                              self.checkerboard.ptr, self.index.ptr))

    def reconstruct_upscale(self, common, depth: DeviceBuffer, hdr: DeviceBuffer):
        """reconstruct_pass + upscale in one launch."""
        w, h = self.viewport
        previous = self.reconstruct[1].ptr if self.frames > 0 else None
        _abi.check(_abi.lib().skyhip_cloud_reconstruct_upscale(
            self.ctx.handle, C.byref(common), self.checkerboard.ptr, self.index.ptr, self.render_texture.ptr,
            self.distance.ptr, previous, self.reconstruct[0].ptr, depth.ptr, hdr.ptr, w, h))

    def render(self, atmosphere_block, common, cloud, material: _abi.CloudMaterial, transmittance: DeviceBuffer,
               t_size, ap_l: DeviceBuffer, ap_t: DeviceBuffer, ap_size, blue_noise: DeviceBuffer, blue_noise_size,
               depth: DeviceBuffer, hdr: DeviceBuffer, work_counters: DeviceBuffer | None = None, fused=True):
        """One frame of VolumetricCloud::Render: hdr (RGBA16F viewport) is composited in place.
        fused: the checkerboard + index and reconstruction + upscale passes run as one launch
        each (same results bit for bit); otherwise one launch per reference pass."""
        w, h = self.viewport
        if fused:
            self.checkerboard_index(common, depth)
        else:
            self.checkerboard_depth(depth)
            self.shading_index(common)
        cloud_raymarch(self.ctx, atmosphere_block, common, cloud, material, transmittance, t_size, ap_l, ap_t, ap_size,
                       blue_noise, blue_noise_size, self.render_texture, self.distance, w // 4, h // 4,
                       depth=self.checkerboard, shading_index=self.index, work_counters=work_counters)
        if fused:
            self.reconstruct_upscale(common, depth, hdr)
        else:
            self.reconstruct_pass(common)
            self.upscale(common, depth, hdr)
        self.reconstruct.reverse()  # :419-421
        self.frames += 1

    Render = render
