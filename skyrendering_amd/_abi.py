"""ctypes view of include/skyhip.h: the std140 parameter blocks and the C entry points.

The library is loaded from this package directory (libskyhip.so, built in-tree by
__graft_entry__.build()).  There is no fallback: if the library is missing, importing
the compute entry points raises, so no caller can silently run without the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libskyhip.so")

f32 = C.c_float
i32 = C.c_int32
u32 = C.c_uint32
vp = C.c_void_p


def _arr(t, n):
    return t * n


class SkyhipError(RuntimeError):
    """A skyhip call returned a negative status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"skyhip error {code}: {message}")
        self.code = code


OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_INVALID_SIZE = -2
ERR_NON_FINITE = -3
ERR_DEVICE = -4
ERR_OUT_OF_MEMORY = -5
ERR_NOT_READY = -6


class AtmosphereBufferData(C.Structure):
    _fields_ = [
        ("solar_illuminance", _arr(f32, 3)), ("sun_angular_radius", f32),
        ("rayleigh_scattering", _arr(f32, 3)), ("inv_rayleigh_exponential_distribution", f32),
        ("mie_scattering", _arr(f32, 3)), ("inv_mie_exponential_distribution", f32),
        ("mie_absorption", _arr(f32, 3)), ("ozone_center_altitude", f32),
        ("ozone_absorption", _arr(f32, 3)), ("inv_ozone_width", f32),
        ("ground_albedo", _arr(f32, 3)), ("mie_phase_g", f32),
        ("_atmosphere_padding", _arr(f32, 3)), ("multiscattering_mask", f32),
        ("bottom_radius", f32), ("top_radius", f32), ("transmittance_steps", f32),
        ("multiscattering_steps", f32),
    ]


class AtmosphereParametersC(C.Structure):
    _fields_ = [
        ("solar_illuminance", _arr(f32, 3)), ("sun_angular_radius", f32),
        ("bottom_radius", f32), ("thickness", f32), ("ground_albedo", _arr(f32, 3)),
        ("rayleigh_exponential_distribution", f32), ("rayleigh_scattering_scale", f32),
        ("rayleigh_scattering", _arr(f32, 3)),
        ("mie_exponential_distribution", f32), ("mie_phase_g", f32), ("mie_scattering_scale", f32),
        ("mie_scattering", _arr(f32, 3)), ("mie_absorption_scale", f32), ("mie_absorption", _arr(f32, 3)),
        ("ozone_center_altitude", f32), ("ozone_width", f32), ("ozone_absorption_scale", f32),
        ("ozone_absorption", _arr(f32, 3)),
        ("transmittance_steps", f32), ("multiscattering_steps", f32), ("multiscattering_mask", f32),
    ]


class AtmosphereRenderBufferData(C.Structure):
    _fields_ = [
        ("sun_direction", _arr(f32, 3)), ("star_luminance_scale", f32),
        ("earth_center", _arr(f32, 3)), ("camera_earth_center_distance", f32),
        ("camera_position", _arr(f32, 3)), ("raymarching_steps", f32),
        ("up_direction", _arr(f32, 3)), ("sky_view_lut_steps", f32),
        ("right_direction", _arr(f32, 3)), ("aerial_perspective_lut_steps", f32),
        ("front_direction", _arr(f32, 3)), ("aerial_perspective_lut_max_distance", f32),
        ("moon_position", _arr(f32, 3)), ("moon_radius", f32),
        ("inv_view_projection", _arr(f32, 16)), ("light_view_projection", _arr(f32, 16)),
        ("padding00", f32), ("uInvShadowFroxelMaxDistance", f32), ("blocker_kernel_size_k", f32),
        ("pcss_size_k", f32), ("uCloudShadowMapMat", _arr(f32, 16)),
    ]


class AtmosphereRenderParametersC(C.Structure):
    _fields_ = [
        ("sun_direction_theta", f32), ("sun_direction_phi", f32), ("star_luminance_scale", f32),
        ("raymarching_steps", f32), ("sky_view_lut_steps", f32), ("aerial_perspective_lut_steps", f32),
        ("aerial_perspective_lut_max_distance", f32), ("camera_position", _arr(f32, 3)),
        ("view_projection", _arr(f32, 16)),
    ]


class CameraC(C.Structure):
    _fields_ = [
        ("position", _arr(f32, 3)), ("yaw", f32), ("pitch", f32), ("aspect", f32), ("fovy", f32),
        ("z_near", f32), ("z_far", f32),
    ]


class CloudCommonBufferData(C.Structure):
    _fields_ = [
        ("uInvMVP", _arr(f32, 16)), ("uReprojectMat", _arr(f32, 16)), ("uLightVP", _arr(f32, 16)),
        ("uInvLightVP", _arr(f32, 16)), ("uShadowMapReprojectMat", _arr(f32, 16)),
        ("uCameraPos", _arr(f32, 3)), ("uBaseShadingIndex", u32),
        ("uLinearDepthParam", _arr(f32, 2)), ("uBottomAltitude", f32), ("uTopAltitude", f32),
        ("uSunDirection", _arr(f32, 3)), ("uFrameID", f32),
        ("uInvShadowFroxelMaxDistance", f32), ("uAerialPerspectiveLutMaxDistance", f32),
        ("uShadowFroxelMaxDistance", f32), ("uEarthRadius", f32),
    ]


class CloudBufferData(C.Structure):
    _fields_ = [
        ("uSunIlluminanceScale", f32), ("uMaxRaymarchDistance", f32), ("uMaxRaymarchSteps", f32),
        ("uMaxVisibleDistance", f32), ("uEnvColorScale", _arr(f32, 3)), ("uShadowSteps", f32),
        ("uSunMultiscatteringSigmaScale", f32), ("uEnvMultiscatteringSigmaScale", f32),
        ("uShadowDistance", f32), ("uEnvBottomVisibility", f32), ("padding_", _arr(f32, 3)),
        ("uEnvSunHeightCurveExp", f32),
    ]


class VolumetricCloudParametersC(C.Structure):
    _fields_ = [
        ("bottom_altitude", f32), ("thickness", f32), ("max_raymarch_distance", f32),
        ("max_raymarch_steps", f32), ("max_visible_distance", f32), ("env_color", _arr(f32, 3)),
        ("env_color_scale", f32), ("sun_illuminance_scale", f32), ("shadow_steps", f32),
        ("shadow_distance", f32), ("shadow_map_max_distance", f32), ("shadow_froxel_max_distance", f32),
        ("sun_multiscattering_sigma_scale", f32), ("env_multiscattering_sigma_scale", f32),
        ("env_bottom_visibility", f32), ("env_sun_height_curve_exp", f32),
    ]


class SampleInfo(C.Structure):
    _fields_ = [("bias", _arr(f32, 2)), ("frequency", f32), ("k_lod", f32)]


class CloudMaterialCommonBufferData(C.Structure):
    _fields_ = [
        ("uCloudMapSampleInfo", SampleInfo), ("uDetailSampleInfo", SampleInfo),
        ("uDisplacementSampleInfo", SampleInfo), ("padding0", _arr(f32, 2)), ("uLodBias", f32),
        ("uDensity", f32),
    ]


class CloudMaterial1BufferData(C.Structure):
    _fields_ = [
        ("uBaseDensityThreshold", f32), ("uBaseHeightHardness", f32), ("uBaseEdgeHardness", f32),
        ("uDetailBase", f32), ("uDetailScale", f32), ("uHeightCut", f32), ("uEdgeCur", f32),
        ("padding1", f32),
    ]


class CloudMaterial0BufferData(C.Structure):
    _fields_ = [("uDetailParam", _arr(f32, 2)), ("uDisplacementScale", f32), ("padding1", f32)]


class CloudMinimalBufferData(C.Structure):
    _fields_ = [("padding", _arr(f32, 3)), ("uDensity", f32)]


class CloudVoxelBufferData(C.Structure):
    _fields_ = [
        ("uSampleFrequency", _arr(f32, 2)), ("uLodBias", f32), ("uDensity", f32),
        ("uSampleBias", _arr(f32, 2)), ("uSampleLodK", f32), ("voxel_material_padding", f32),
    ]


MATERIAL_DEFAULT0, MATERIAL_DEFAULT1, MATERIAL_MINIMAL, MATERIAL_VOXEL, MATERIAL_VOXEL_F16 = 0, 1, 2, 3, 4
MATERIAL_DEFAULT1_WORLEY = 5


class CloudWorleyDetailBufferData(C.Structure):
    _fields_ = [("uWorleySampleInfo", SampleInfo), ("uWorleyBase", f32), ("uWorleyScale", f32),
                ("padding", _arr(f32, 2))]


class CloudPhase(C.Structure):
    _fields_ = [("back_g", f32), ("back_scale", f32), ("forward_g", f32), ("padding", f32)]


class CloudMaterial(C.Structure):
    _fields_ = [
        ("kind", i32), ("common", CloudMaterialCommonBufferData), ("material1", CloudMaterial1BufferData),
        ("material0", CloudMaterial0BufferData), ("minimal", CloudMinimalBufferData),
        ("voxel", CloudVoxelBufferData),
        ("cloud_map", vp), ("cloud_map_size", i32), ("cloud_map_levels", i32),
        ("detail", vp), ("detail_size", i32), ("detail_levels", i32),
        ("displacement", vp), ("displacement_size", i32), ("displacement_levels", i32),
        ("voxel_texture", vp), ("voxel_dim", _arr(i32, 3)), ("voxel_levels", i32),
        ("voxel_grid", vp), ("voxel_grid_dim", _arr(i32, 3)), ("voxel_grid_padding", i32),
        ("worley", CloudWorleyDetailBufferData), ("worley_texture", vp), ("worley_size", i32),
        ("worley_levels", i32), ("cloud_map_cells", vp), ("detail_cells", vp), ("worley_cells", vp),
    ]


class NoiseCreateInfo(C.Structure):
    _fields_ = [("seed", u32), ("base_frequency", u32), ("remap_min", f32), ("remap_max", f32)]


PRNG_WANG_HASH, PRNG_PCG_HASH = 0, 1
ENV_OFF, ENV_CONST_ENVIRONMENT_MAP, ENV_GROUND_SINGLE_BOUNCE, ENV_GROUND_MULTI_BOUNCE = 0, 1, 2, 3


class PathTracingParametersC(C.Structure):
    _fields_ = [
        ("max_bounces", i32), ("region_box_half_width", f32), ("importance_sampling", i32),
        ("forward_phase_g", f32), ("back_phase_g", f32), ("forward_scattering_ratio", f32), ("prng", i32),
        ("environment_lighting", i32), ("sigma_t_max", f32), ("albedo", f32), ("model3", _arr(f32, 9)),
    ]


class PathTracingInputsC(C.Structure):
    _fields_ = [
        ("transmittance", vp), ("transmittance_width", i32), ("transmittance_height", i32),
        ("sky_view_luminance", vp), ("sky_view_transmittance", vp), ("sky_view_width", i32),
        ("sky_view_height", i32), ("ap_luminance", vp), ("ap_transmittance", vp), ("ap_width", i32),
        ("ap_height", i32), ("ap_depth", i32), ("brick_size", i32), ("brick_majorants", vp), ("cell_corners", vp),
    ]


class VoxelGridCreateInfoC(C.Structure):
    _fields_ = [
        ("seed", u32), ("base_frequency", u32), ("octaves", i32), ("fbm_inv_std", f32), ("lognormal_sigma", f32),
        ("threshold", f32), ("ellipsoid_center", _arr(f32, 3)), ("ellipsoid_radii", _arr(f32, 3)),
    ]


class LutBatchSetC(C.Structure):
    _fields_ = [("atmosphere", AtmosphereBufferData), ("transmittance_width", i32), ("transmittance_height", i32),
                ("multiscattering_width", i32), ("multiscattering_height", i32)]


class LutBatchViewC(C.Structure):
    _fields_ = [("set", i32), ("sky_view_width", i32), ("sky_view_height", i32), ("padding", i32),
                ("render", AtmosphereRenderBufferData)]


class LutBatchTuningC(C.Structure):
    _fields_ = [(n, i32) for n in (
        "transmittance_block", "transmittance_tile", "transmittance_blocks_per_cu", "multiscattering_waves",
        "multiscattering_blocks_per_cu", "sky_view_block", "sky_view_tile", "sky_view_blocks_per_cu",
        "sky_view_largest_first")]


assert C.sizeof(AtmosphereBufferData) == 128
assert C.sizeof(LutBatchSetC) == 144
assert C.sizeof(LutBatchViewC) == 336
assert C.sizeof(LutBatchTuningC) == 36
assert C.sizeof(AtmosphereRenderBufferData) == 320
assert C.sizeof(CloudCommonBufferData) == 384
assert C.sizeof(CloudBufferData) == 64
assert C.sizeof(CloudMaterialCommonBufferData) == 64
assert C.sizeof(CloudMaterial1BufferData) == 32
assert C.sizeof(CloudVoxelBufferData) == 32
assert C.sizeof(PathTracingParametersC) == 76
assert C.sizeof(PathTracingInputsC) == 88
assert C.sizeof(VoxelGridCreateInfoC) == 48

P = C.POINTER
_PROTOTYPES = {
    "skyhip_create": (i32, [C.c_int, P(vp)]),
    "skyhip_destroy": (i32, [vp]),
    "skyhip_last_error": (C.c_char_p, []),
    "skyhip_stream": (vp, [vp]),
    "skyhip_synchronize": (i32, [vp]),
    "skyhip_malloc": (i32, [vp, C.c_size_t, P(vp)]),
    "skyhip_free": (i32, [vp, vp]),
    "skyhip_memcpy_h2d": (i32, [vp, vp, vp, C.c_size_t]),
    "skyhip_memcpy_d2h": (i32, [vp, vp, vp, C.c_size_t]),
    "skyhip_memcpy_d2d": (i32, [vp, vp, vp, C.c_size_t]),
    "skyhip_memset": (i32, [vp, vp, C.c_int, C.c_size_t]),
    "skyhip_event_record": (i32, [vp, C.c_int]),
    "skyhip_event_elapsed_ms": (i32, [vp, C.c_int, C.c_int, P(f32)]),
    "skyhip_device_info": (i32, [vp, P(i32), P(i32), P(i32), C.c_char_p, i32]),
    "skyhip_atmosphere_assign_buffer_data": (i32, [P(AtmosphereParametersC), P(AtmosphereBufferData)]),
    "skyhip_atmosphere_render_assign_buffer_data": (
        i32, [P(AtmosphereRenderParametersC), f32, P(AtmosphereRenderBufferData)]),
    "skyhip_camera_view_projection": (i32, [P(CameraC), P(f32), P(f32), P(f32), P(f32)]),
    "skyhip_cloud_assign_buffer_data": (
        i32, [P(VolumetricCloudParametersC), P(CameraC), f32, P(f32), f32, i32, P(CloudCommonBufferData),
              P(CloudBufferData), P(f32)]),
    "skyhip_cloud_sample_info": (i32, [f32, i32, f32, f32, f32, C.c_double, C.c_double, P(SampleInfo)]),
    "skyhip_transmittance_lut": (i32, [vp, P(AtmosphereBufferData), vp, i32, i32]),
    "skyhip_multiscattering_lut": (i32, [vp, P(AtmosphereBufferData), vp, i32, i32, vp, i32, i32]),
    "skyhip_atmosphere_update_luts": (i32, [vp, P(AtmosphereParametersC), vp, i32, i32, vp, i32, i32]),
    "skyhip_sky_view_lut": (
        i32, [vp, P(AtmosphereBufferData), P(AtmosphereRenderBufferData), vp, i32, i32, vp, i32, i32, vp, i32,
              vp, vp, i32, i32]),
    "skyhip_aerial_perspective_lut": (
        i32, [vp, P(AtmosphereBufferData), P(AtmosphereRenderBufferData), vp, i32, i32, vp, i32, i32, vp, i32,
              vp, vp, i32, i32, i32]),
}

_lib = None


def lib():
    """Loads libskyhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Names of every function include/skyhip.h declares (parsed from the header)."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "skyhip.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:int|void\*|const char\*|size_t|long long)\s+(skyhip_\w+)\s*\(", text, re.M)))


def register(name, restype, argtypes):
    """Adds a prototype (used by modules that bind further entry points)."""
    _PROTOTYPES[name] = (restype, argtypes)
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


def check(rc: int):
    if rc != OK:
        msg = lib().skyhip_last_error()
        raise SkyhipError(rc, msg.decode() if msg else "")
    return rc
