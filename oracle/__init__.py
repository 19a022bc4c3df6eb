"""The oracle (test infrastructure): ctypes bindings of oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this, as the
checker or as the timed CPU baseline; the product package never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from skyrendering_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

P = C.POINTER
vp = C.c_void_p
i32 = C.c_int
f32 = C.c_float
f64 = C.c_double

_PROTO = {
    "oracle_max_threads": (i32, []),
    "oracle_set_threads": (None, [i32]),
    "oracle_transmittance_lut_f64": (None, [P(_abi.AtmosphereBufferData), i32, i32, vp]),
    "oracle_transmittance_lut_f32": (None, [P(_abi.AtmosphereBufferData), i32, i32, vp]),
    "oracle_multiscattering_lut_f64": (None, [P(_abi.AtmosphereBufferData), vp, i32, i32, i32, i32, vp]),
    "oracle_multiscattering_lut_f32": (None, [P(_abi.AtmosphereBufferData), vp, i32, i32, i32, i32, vp]),
    "oracle_sky_view_lut_f64": (None, [P(_abi.AtmosphereBufferData), P(_abi.AtmosphereRenderBufferData), vp, i32,
                                       i32, vp, i32, i32, vp, i32, i32, i32, vp, vp]),
    "oracle_sky_view_lut_f32": (None, [P(_abi.AtmosphereBufferData), P(_abi.AtmosphereRenderBufferData), vp, i32,
                                       i32, vp, i32, i32, vp, i32, i32, i32, vp, vp]),
    "oracle_aerial_perspective_lut_f64": (None, [P(_abi.AtmosphereBufferData), P(_abi.AtmosphereRenderBufferData),
                                                 vp, i32, i32, vp, i32, i32, vp, i32, i32, i32, i32, vp, vp]),
    "oracle_aerial_perspective_lut_f32": (None, [P(_abi.AtmosphereBufferData), P(_abi.AtmosphereRenderBufferData),
                                                 vp, i32, i32, vp, i32, i32, vp, i32, i32, i32, i32, vp, vp]),
    "oracle_extinction_f64": (f64, [P(_abi.AtmosphereBufferData), f64, i32]),
    "oracle_mie_phase_f64": (f64, [f64, f64]),
    "oracle_rayleigh_phase_f64": (f64, [f64]),
    "oracle_transmittance_rmu_f64": (None, [P(_abi.AtmosphereBufferData), f64, f64, vp]),
    "oracle_transmittance_index_to_rmu_f64": (None, [P(_abi.AtmosphereBufferData), i32, i32, i32, i32, P(f64),
                                                     P(f64)]),
    "oracle_transmittance_rmu_to_uv_f64": (None, [P(_abi.AtmosphereBufferData), i32, i32, f64, f64, P(f64),
                                                  P(f64)]),
    "oracle_sky_view_index_roundtrip_f64": (None, [P(_abi.AtmosphereBufferData), f64, i32, i32, i32, i32, P(f64),
                                                   P(f64)]),
    "oracle_scattered_luminance_f64": (None, [P(_abi.AtmosphereBufferData), vp, i32, i32, vp, i32, i32, vp, vp, vp,
                                              f64, f32, f32, vp, vp]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for k, (r, a) in _PROTO.items():
            fn = getattr(L, k)
            fn.restype = r
            fn.argtypes = a
        _lib = L
    return _lib


def register(name, restype, argtypes):
    _PROTO[name] = (restype, argtypes)
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


def _ptr(a):
    return None if a is None else a.ctypes.data


def _dt(precision):
    return np.float64 if precision == "f64" else np.float32


def transmittance_lut(atm, w, h, precision="f64"):
    out = np.empty((h, w, 4), dtype=_dt(precision))
    getattr(lib(), f"oracle_transmittance_lut_{precision}")(C.byref(atm), w, h, _ptr(out))
    return out


def multiscattering_lut(atm, T, w, h, precision="f64"):
    T = np.ascontiguousarray(T, dtype=np.float32)
    out = np.empty((h, w, 4), dtype=_dt(precision))
    getattr(lib(), f"oracle_multiscattering_lut_{precision}")(C.byref(atm), _ptr(T), T.shape[1], T.shape[0], w, h,
                                                              _ptr(out))
    return out


def sky_view_lut(atm, render, T, MS, w, h, blue_noise=None, precision="f64"):
    T = np.ascontiguousarray(T, dtype=np.float32)
    MS = np.ascontiguousarray(MS, dtype=np.float32)
    bn = None if blue_noise is None else np.ascontiguousarray(blue_noise, dtype=np.uint16)
    L = np.empty((h, w, 4), dtype=_dt(precision))
    Tr = np.empty((h, w, 4), dtype=_dt(precision))
    getattr(lib(), f"oracle_sky_view_lut_{precision}")(
        C.byref(atm), C.byref(render), _ptr(T), T.shape[1], T.shape[0], _ptr(MS), MS.shape[1], MS.shape[0], _ptr(bn),
        0 if bn is None else bn.shape[0], w, h, _ptr(L), _ptr(Tr))
    return L, Tr


def aerial_perspective_lut(atm, render, T, MS, w, h, d, blue_noise=None, precision="f64"):
    T = np.ascontiguousarray(T, dtype=np.float32)
    MS = np.ascontiguousarray(MS, dtype=np.float32)
    bn = None if blue_noise is None else np.ascontiguousarray(blue_noise, dtype=np.uint16)
    L = np.empty((d, h, w, 4), dtype=_dt(precision))
    Tr = np.empty((d, h, w, 4), dtype=_dt(precision))
    getattr(lib(), f"oracle_aerial_perspective_lut_{precision}")(
        C.byref(atm), C.byref(render), _ptr(T), T.shape[1], T.shape[0], _ptr(MS), MS.shape[1], MS.shape[0], _ptr(bn),
        0 if bn is None else bn.shape[0], w, h, d, _ptr(L), _ptr(Tr))
    return L, Tr


def max_threads():
    return lib().oracle_max_threads()


def set_threads(n):
    lib().oracle_set_threads(int(n))


# ---------------------------------------------------------------------------
# Noise and cloud raymarch
# ---------------------------------------------------------------------------
class CloudImagesC(C.Structure):
    _fields_ = [("transmittance", vp), ("tw", i32), ("th", i32), ("ap_luminance", vp), ("ap_transmittance", vp),
                ("aw", i32), ("ah", i32), ("ad", i32), ("blue_noise", vp), ("bn_size", i32), ("depth", vp),
                ("index", vp), ("image_w", i32), ("image_h", i32)]


_NI = P(_abi.NoiseCreateInfo)
_PROTO.update({
    "oracle_cloud_map_texels": (None, [_NI, _NI, i32, vp, C.c_long, vp, i32]),
    "oracle_detail_texels": (None, [_NI, _NI, i32, vp, C.c_long, vp, i32]),
    "oracle_displacement_texels": (None, [_NI, i32, vp, C.c_long, vp, i32]),
    "oracle_worley_texels": (None, [_NI, i32, vp, C.c_long, vp, i32]),
    "oracle_perlin_noise": (f64, [f64, f64, f64, C.c_uint, C.c_uint]),
    "oracle_worley_noise": (f64, [f64, f64, f64, C.c_uint, C.c_uint]),
    "oracle_wang_hash": (C.c_uint, [C.c_uint]),
    "oracle_pcg_hash": (C.c_uint, [C.c_uint]),
    "oracle_generate_mipmaps": (None, [vp, i32, i32, i32, i32, i32, i32]),
    "oracle_cloud_raymarch_f64": (None, [P(_abi.AtmosphereBufferData), P(_abi.CloudCommonBufferData),
                                         P(_abi.CloudBufferData), P(_abi.CloudMaterial), P(CloudImagesC), i32, vp,
                                         C.c_long, vp, vp, vp]),
    "oracle_cloud_raymarch_f32": (None, [P(_abi.AtmosphereBufferData), P(_abi.CloudCommonBufferData),
                                         P(_abi.CloudBufferData), P(_abi.CloudMaterial), P(CloudImagesC), i32, vp,
                                         C.c_long, vp, vp, vp]),
    "oracle_cloud_raymarch_forced_f64": (None, [P(_abi.AtmosphereBufferData), P(_abi.CloudCommonBufferData),
                                                P(_abi.CloudBufferData), P(_abi.CloudMaterial), P(CloudImagesC), i32,
                                                vp, C.c_long, vp, vp, vp, vp, vp, vp, vp, vp]),
    "oracle_cloud_raymarch_forced_f32": (None, [P(_abi.AtmosphereBufferData), P(_abi.CloudCommonBufferData),
                                                P(_abi.CloudBufferData), P(_abi.CloudMaterial), P(CloudImagesC), i32,
                                                vp, C.c_long, vp, vp, vp, vp, vp, vp, vp, vp]),
    "oracle_cloud_trace_f64": (i32, [P(_abi.AtmosphereBufferData), P(_abi.CloudCommonBufferData),
                                     P(_abi.CloudBufferData), P(_abi.CloudMaterial), P(CloudImagesC), i32, i32, vp,
                                     i32]),
    "oracle_cloud_sigma_t_f64": (f64, [P(_abi.CloudCommonBufferData), P(_abi.CloudBufferData),
                                       P(_abi.CloudMaterial), f64, f64, f64]),
    "oracle_cloud_shadow_f64": (f64, [P(_abi.CloudCommonBufferData), P(_abi.CloudBufferData),
                                      P(_abi.CloudMaterial), f64, f64, f64]),
})


def cloud_map_texels(density, height, size, texels, precision="f64"):
    t = np.ascontiguousarray(texels, dtype=np.int32)
    out = np.empty((len(t), 2), np.uint8)
    lib().oracle_cloud_map_texels(C.byref(density), C.byref(height), size, _ptr(t), len(t), _ptr(out),
                                  1 if precision == "f64" else 0)
    return out


def detail_texels(perlin, worley, size, texels, precision="f64"):
    t = np.ascontiguousarray(texels, dtype=np.int32)
    out = np.empty(len(t), np.uint8)
    lib().oracle_detail_texels(C.byref(perlin), C.byref(worley), size, _ptr(t), len(t), _ptr(out),
                               1 if precision == "f64" else 0)
    return out


def displacement_texels(perlin, size, texels, precision="f64"):
    t = np.ascontiguousarray(texels, dtype=np.int32)
    out = np.empty((len(t), 4), np.uint8)
    lib().oracle_displacement_texels(C.byref(perlin), size, _ptr(t), len(t), _ptr(out),
                                     1 if precision == "f64" else 0)
    return out


def worley_texels(worley, size, texels, precision="f64"):
    t = np.ascontiguousarray(texels, dtype=np.int32)
    out = np.empty(len(t), np.uint8)
    lib().oracle_worley_texels(C.byref(worley), size, _ptr(t), len(t), _ptr(out), 1 if precision == "f64" else 0)
    return out


def generate_mipmaps(chain: np.ndarray, dims, channels, w, h, d, levels):
    chain = np.ascontiguousarray(chain, dtype=np.uint8)
    lib().oracle_generate_mipmaps(_ptr(chain), dims, channels, w, h, d, levels)
    return chain


class HostCloudInputs:
    """Host copies of every raymarch input, kept alive for the oracle call."""

    def __init__(self, material, cloud_map=None, detail=None, displacement=None, voxel=None, T=None, ap_l=None,
                 ap_t=None, blue_noise=None, depth=None, index=None, image_size=None, worley=None):
        self.keep = [x for x in (cloud_map, detail, displacement, voxel, T, ap_l, ap_t, blue_noise, depth, index,
                                 worley) if x is not None]
        m = _abi.CloudMaterial()
        C.memmove(C.byref(m), C.byref(material), C.sizeof(m))
        m.cloud_map = _ptr(cloud_map)
        m.detail = _ptr(detail)
        m.displacement = _ptr(displacement)
        m.voxel_texture = _ptr(voxel)
        m.worley_texture = _ptr(worley)
        self.material = m
        im = CloudImagesC()
        im.transmittance, im.tw, im.th = _ptr(T), T.shape[1], T.shape[0]
        im.ap_luminance, im.ap_transmittance = _ptr(ap_l), _ptr(ap_t)
        im.ad, im.ah, im.aw = ap_l.shape[0], ap_l.shape[1], ap_l.shape[2]
        im.blue_noise, im.bn_size = _ptr(blue_noise), blue_noise.shape[0]
        im.depth, im.index = _ptr(depth), _ptr(index)
        im.image_w, im.image_h = image_size
        self.images = im


def cloud_raymarch(atm, common, cloud, inputs: HostCloudInputs, width, pixels, precision="f64"):
    px = np.ascontiguousarray(pixels, dtype=np.int32)
    dt = _dt(precision)
    rgba = np.empty((len(px), 4), dt)
    dist = np.empty(len(px), dt)
    steps = np.empty(len(px), np.int32)
    getattr(lib(), f"oracle_cloud_raymarch_{precision}")(
        C.byref(atm), C.byref(common), C.byref(cloud), C.byref(inputs.material), C.byref(inputs.images), width,
        _ptr(px), len(px), _ptr(rgba), _ptr(dist), _ptr(steps))
    return rgba, dist, steps


def cloud_raymarch_forced(atm, common, cloud, inputs: HostCloudInputs, width, pixels, forced_steps, precision="f64"):
    """Checker mode: replays another implementation's per-pixel step decisions
    (uint32[n, 2]: executed, scheduled; each segment 0 | segment 1 << 16) and returns
    (rgba, distance, steps, early-exit margin, step-count margin, texel margin, LOD margin): the
    last two are each pixel's smallest distance of a NEAREST texel coordinate from an integer
    (texel units) and of a level-of-detail from a switch point, over every fetch of its march.
    precision "f32" runs the oracle's FP32 instantiation under the same decisions."""
    px = np.ascontiguousarray(pixels, dtype=np.int32)
    fs = np.ascontiguousarray(forced_steps, dtype=np.uint32).reshape(-1, 2)
    dt = _dt(precision)
    rgba = np.empty((len(px), 4), dt)
    dist = np.empty(len(px), dt)
    steps = np.empty(len(px), np.int32)
    margin = np.empty(len(px), np.float64)
    step_margin = np.empty(len(px), np.float64)
    texel_margin = np.empty(len(px), np.float64)
    lod_margin = np.empty(len(px), np.float64)
    getattr(lib(), f"oracle_cloud_raymarch_forced_{precision}")(
        C.byref(atm), C.byref(common), C.byref(cloud), C.byref(inputs.material), C.byref(inputs.images), width,
        _ptr(px), len(px), _ptr(fs), _ptr(rgba), _ptr(dist), _ptr(steps), _ptr(margin), _ptr(step_margin),
        _ptr(texel_margin), _ptr(lod_margin))
    return rgba, dist, steps, margin, step_margin, texel_margin, lod_margin


# ---------------------------------------------------------------------------
# Voxel-cloud path tracer
# ---------------------------------------------------------------------------
_PT_ARGS = [P(_abi.AtmosphereBufferData), P(_abi.AtmosphereRenderBufferData), P(_abi.CloudCommonBufferData),
            P(_abi.CloudMaterial), P(_abi.PathTracingParametersC), P(_abi.PathTracingInputsC), i32, i32, vp, C.c_long,
            C.c_uint32, i32, vp]
register("oracle_path_trace_f64", None, _PT_ARGS)
register("oracle_path_trace_f32", None, _PT_ARGS)
register("oracle_pt_hg_sample_f64", None, [P(_abi.PathTracingParametersC), C.c_uint32, vp, vp, P(f64),
                                           P(C.c_uint32)])
register("oracle_pt_transmittance_f64", f64, [P(_abi.CloudCommonBufferData), P(_abi.CloudMaterial),
                                              P(_abi.PathTracingParametersC), C.c_uint32, vp, vp])
register("oracle_pt_track_f64", f64, [P(_abi.CloudCommonBufferData), P(_abi.CloudMaterial),
                                      P(_abi.PathTracingParametersC), vp, i32, i32, C.c_uint32, vp, vp])
register("oracle_pt_sigma_t_f64", f64, [P(_abi.CloudCommonBufferData), P(_abi.CloudMaterial),
                                        P(_abi.PathTracingParametersC), f64, f64, f64])
register("oracle_voxel_grid_f16", None, [P(_abi.VoxelGridCreateInfoC), i32, i32, i32, vp])
register("oracle_voxel_grid_density_f32", None, [P(_abi.VoxelGridCreateInfoC), i32, i32, i32, vp, C.c_long, vp])
register("oracle_brick_majorants", None, [vp, i32, i32, i32, i32, vp])
register("oracle_float_to_half", C.c_uint16, [f32])
register("oracle_half_to_float", f32, [C.c_uint16])


class HostPathTraceInputs:
    """Host copies of the path tracer's inputs, kept alive for the oracle call.
    grid: uint16 FP16 bits [z, y, x] (VOXEL_F16 material), T / sv_l / sv_t: RGBA32F 2D LUTs,
    ap_l / ap_t: RGBA32F [d, h, w, 4]; majorants: float32 [bz, by, bx] brick majorants (tracking
    over the brick grid instead of the single majorant), built with brick edge `brick`."""

    def __init__(self, material, T, sv_l, sv_t, ap_l, ap_t, grid=None, cloud_map=None, detail=None,
                 displacement=None, voxel=None, majorants=None, brick=None):
        if majorants is not None:
            majorants = np.asarray(majorants, np.float32)
        arrs = [np.ascontiguousarray(x) if x is not None else None
                for x in (T, sv_l, sv_t, ap_l, ap_t, grid, cloud_map, detail, displacement, voxel, majorants)]
        T, sv_l, sv_t, ap_l, ap_t, grid, cloud_map, detail, displacement, voxel, majorants = arrs
        self.keep = [x for x in arrs if x is not None]
        m = _abi.CloudMaterial()
        C.memmove(C.byref(m), C.byref(material), C.sizeof(m))
        m.cloud_map = _ptr(cloud_map)
        m.detail = _ptr(detail)
        m.displacement = _ptr(displacement)
        m.voxel_texture = _ptr(voxel)
        m.voxel_grid = _ptr(grid)
        self.material = m
        i = _abi.PathTracingInputsC()
        i.transmittance, i.transmittance_width, i.transmittance_height = _ptr(T), T.shape[1], T.shape[0]
        i.sky_view_luminance, i.sky_view_transmittance = _ptr(sv_l), _ptr(sv_t)
        i.sky_view_width, i.sky_view_height = sv_l.shape[1], sv_l.shape[0]
        i.ap_luminance, i.ap_transmittance = _ptr(ap_l), _ptr(ap_t)
        i.ap_depth, i.ap_height, i.ap_width = ap_l.shape[0], ap_l.shape[1], ap_l.shape[2]
        i.brick_majorants = _ptr(majorants)
        i.brick_size = 0 if brick is None else int(brick)
        self.inputs = i


def path_trace(atm, render, common, params, inputs: HostPathTraceInputs, width, height, pixels, first_frame, frames,
               precision="f64"):
    """Samples of the tracer (single majorant, or brick majorants when the inputs hold them):
    array [len(pixels), frames, 4]."""
    px = np.ascontiguousarray(pixels, dtype=np.int32)
    out = np.empty((len(px), frames, 4), _dt(precision))
    getattr(lib(), f"oracle_path_trace_{precision}")(
        C.byref(atm), C.byref(render), C.byref(common), C.byref(inputs.material), C.byref(params),
        C.byref(inputs.inputs), width, height, _ptr(px), len(px), first_frame, frames, _ptr(out))
    return out


def pt_hg_sample(params, seed, d):
    d = np.ascontiguousarray(d, dtype=np.float64)
    out = np.empty(3, np.float64)
    vop = C.c_double()
    s = C.c_uint32()
    lib().oracle_pt_hg_sample_f64(C.byref(params), seed, _ptr(d), _ptr(out), C.byref(vop), C.byref(s))
    return out, vop.value, s.value


def pt_transmittance(common, material, params, seed, o, d):
    o = np.ascontiguousarray(o, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    return lib().oracle_pt_transmittance_f64(C.byref(common), C.byref(material), C.byref(params), seed, _ptr(o),
                                             _ptr(d))


def pt_track(common, material, params, seed, o, d, mode="delta", majorants=None, brick=None):
    """One tracking call: mode "delta" returns the collision distance (-1: none), "ratio" the
    transmittance estimate; majorants (float32 [bz, by, bx], brick edge `brick`, default
    BRICK_TEXELS) selects brick-majorant tracking."""
    o = np.ascontiguousarray(o, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    m = None if majorants is None else np.ascontiguousarray(majorants, np.float32)
    return lib().oracle_pt_track_f64(C.byref(common), C.byref(material), C.byref(params), _ptr(m),
                                     BRICK_TEXELS if brick is None else int(brick), {"delta": 0, "ratio": 1}[mode],
                                     seed, _ptr(o), _ptr(d))


def pt_sigma_t(common, material, params, p):
    return lib().oracle_pt_sigma_t_f64(C.byref(common), C.byref(material), C.byref(params), *map(float, p))


def voxel_grid_f16(ci, dims):
    """FP16 bits [z, y, x] of the workload grid."""
    dx, dy, dz = dims
    out = np.empty((dz, dy, dx), np.uint16)
    lib().oracle_voxel_grid_f16(C.byref(ci), dx, dy, dz, _ptr(out))
    return out


def voxel_grid_density(ci, dims, xyz):
    xyz = np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1, 3)
    out = np.empty(len(xyz), np.float32)
    lib().oracle_voxel_grid_density_f32(C.byref(ci), *dims, _ptr(xyz), len(xyz), _ptr(out))
    return out


BRICK_TEXELS = 16  # kBrickTexels of oracle_pathtrace.h: the default brick edge


def brick_majorants(grid, brick=None):
    g = np.ascontiguousarray(grid, dtype=np.uint16)
    dz, dy, dx = g.shape
    b = BRICK_TEXELS if brick is None else int(brick)
    out = np.empty(((dz + b - 1) // b, (dy + b - 1) // b, (dx + b - 1) // b), np.float32)
    lib().oracle_brick_majorants(_ptr(g), dx, dy, dz, b, _ptr(out))
    return out


register("oracle_pt_lambert_sample_f64", None, [C.c_uint32, vp, vp, P(C.c_uint32)])


def pt_lambert_sample(seed, n):
    n = np.ascontiguousarray(n, dtype=np.float64)
    out = np.empty(3, np.float64)
    s = C.c_uint32()
    lib().oracle_pt_lambert_sample_f64(seed, _ptr(n), _ptr(out), C.byref(s))
    return out, s.value


register("oracle_sky_radiance_f64", None, [vp, vp, i32, i32, vp, i32, i32, f32, vp, C.c_long, vp])
register("oracle_sky_radiance_f32", None, [vp, vp, i32, i32, vp, i32, i32, f32, vp, C.c_long, vp])


def sky_radiance(atmospheres, T, MS, steps, queries, precision="f64"):
    """Sky radiance of queries [n, 5] (atmosphere, altitude, mu_view, mu_sun, delta_phi);
    atmospheres: a ctypes array of AtmosphereBufferData; T [a, th, tw, 4] / MS [a, mh, mw, 4]
    float32 LUT stacks.  Returns [n, 3] (FP64, or FP32 for precision "f32")."""
    T = np.ascontiguousarray(T, dtype=np.float32)
    MS = np.ascontiguousarray(MS, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.float64)
    out = np.empty((len(q), 3), _dt(precision))
    getattr(lib(), f"oracle_sky_radiance_{precision}")(atmospheres, _ptr(T), T.shape[2], T.shape[1], _ptr(MS),
                                                       MS.shape[2], MS.shape[1], float(steps), _ptr(q), len(q),
                                                       _ptr(out))
    return out


# ---------------------------------------------------------------------------
# Environment-luminance cube and image-based lighting
# ---------------------------------------------------------------------------
_CUBE_ARGS = [P(_abi.AtmosphereBufferData), P(_abi.AtmosphereRenderBufferData), vp, i32, i32, vp, vp, i32, i32, i32, vp]
register("oracle_environment_cube_f64", None, _CUBE_ARGS)
register("oracle_environment_cube_f32", None, _CUBE_ARGS)
register("oracle_cube_mips", None, [vp, i32])
register("oracle_env_radiance_sh_f64", None, [vp, i32, vp])
register("oracle_prefiltered_radiance_f64", None, [vp, i32, i32, i32, vp])
register("oracle_prefiltered_radiance_f32", None, [vp, i32, i32, i32, vp])
register("oracle_cube_sample_f64", None, [vp, i32, i32, vp, f64, vp])
register("oracle_cube_texel_dir_f64", None, [i32, i32, i32, i32, vp])
register("oracle_cube_face_st_f64", None, [vp, P(i32), P(f64), P(f64)])


def cube_texels(width, levels):
    return sum(6 * max(width >> l, 1) ** 2 for l in range(levels))


def cube_levels(width):
    return int(width).bit_length()


def environment_cube(atm, render, T, sv_l, sv_t, width, precision="f64"):
    """The environment-luminance cube with its mip chain: uint16 (binary16 bits) [texels, 4]."""
    T, sv_l, sv_t = (np.ascontiguousarray(x, dtype=np.float32) for x in (T, sv_l, sv_t))
    out = np.empty((cube_texels(width, cube_levels(width)), 4), np.uint16)
    getattr(lib(), f"oracle_environment_cube_{precision}")(
        C.byref(atm), C.byref(render), _ptr(T), T.shape[1], T.shape[0], _ptr(sv_l), _ptr(sv_t), sv_l.shape[1],
        sv_l.shape[0], width, _ptr(out))
    return out


def cube_mips(cube_bits, width):
    """Recomputes levels 1.. of a cube (uint16 [texels, 4]) from its level 0 (2x2 box)."""
    c = np.ascontiguousarray(cube_bits, dtype=np.uint16).copy()
    lib().oracle_cube_mips(_ptr(c), width)
    return c


def env_radiance_sh(cube_bits, width):
    c = np.ascontiguousarray(cube_bits, dtype=np.uint16)
    out = np.empty((9, 4), np.float64)
    lib().oracle_env_radiance_sh_f64(_ptr(c), width, _ptr(out))
    return out


def prefiltered_radiance(cube_bits, width, prefiltered_width, roughness_count, precision="f64"):
    c = np.ascontiguousarray(cube_bits, dtype=np.uint16)
    out = np.empty((cube_texels(prefiltered_width, roughness_count), 4), np.uint16)
    getattr(lib(), f"oracle_prefiltered_radiance_{precision}")(_ptr(c), width, prefiltered_width, roughness_count,
                                                               _ptr(out))
    return out


def cube_sample(cube_bits, width, d, lod):
    c = np.ascontiguousarray(cube_bits, dtype=np.uint16)
    dd = np.ascontiguousarray(d, dtype=np.float64)
    out = np.empty(3, np.float64)
    lib().oracle_cube_sample_f64(_ptr(c), width, cube_levels(width), _ptr(dd), float(lod), _ptr(out))
    return out


def cube_texel_dir(face, x, y, w):
    out = np.empty(3, np.float64)
    lib().oracle_cube_texel_dir_f64(face, x, y, w, _ptr(out))
    return out


def cube_face_st(d):
    dd = np.ascontiguousarray(d, dtype=np.float64)
    f, s, t = C.c_int32(), C.c_double(), C.c_double()
    lib().oracle_cube_face_st_f64(_ptr(dd), C.byref(f), C.byref(s), C.byref(t))
    return f.value, s.value, t.value


# ---------------------------------------------------------------------------
# Cloud shadow map, blur and froxel
# ---------------------------------------------------------------------------
register("oracle_cloud_shadow_map_f64", None, [P(_abi.CloudCommonBufferData), P(_abi.CloudMaterial), vp, i32, vp, i32,
                                               i32, vp])
register("oracle_cloud_shadow_map_f32", None, [P(_abi.CloudCommonBufferData), P(_abi.CloudMaterial), vp, i32, vp, i32,
                                               i32, vp])
register("oracle_cloud_shadow_blur", None, [vp, vp, i32, i32, i32])
register("oracle_cloud_shadow_froxel_f64", None, [P(_abi.CloudCommonBufferData), vp, i32, i32, i32, i32, i32, vp])
register("oracle_cloud_shadow_froxel_f32", None, [P(_abi.CloudCommonBufferData), vp, i32, i32, i32, i32, i32, vp])


def cloud_shadow_map(common, material, blue_noise, w, h, previous=None, precision="f64"):
    """[h, w, 2] (mid-shell depth, transmittance); previous: the previous raw map [h, w, 2] float32."""
    bn = np.ascontiguousarray(blue_noise, dtype=np.uint16)
    prev = None if previous is None else np.ascontiguousarray(previous, dtype=np.float32)
    out = np.empty((h, w, 2), _dt(precision))
    getattr(lib(), f"oracle_cloud_shadow_map_{precision}")(C.byref(common), C.byref(material), _ptr(bn), bn.shape[1],
                                                           _ptr(prev), w, h, _ptr(out))
    return out


def cloud_shadow_blur(rg, axis):
    a = np.ascontiguousarray(rg, dtype=np.float32)
    out = np.empty_like(a)
    lib().oracle_cloud_shadow_blur(_ptr(a), _ptr(out), a.shape[1], a.shape[0], axis)
    return out


def cloud_shadow_froxel(common, shadow_map, w, h, d, precision="f64"):
    m = np.ascontiguousarray(shadow_map, dtype=np.float32)
    out = np.empty((d, h, w), np.uint16)
    getattr(lib(), f"oracle_cloud_shadow_froxel_{precision}")(C.byref(common), _ptr(m), m.shape[1], m.shape[0], w, h,
                                                              d, _ptr(out))
    return out


# Volumetric cloud checkerboard / shading index / reconstruction / upscale.
_cc = P(_abi.CloudCommonBufferData)
register("oracle_cloud_checkerboard_depth", None, [vp, i32, i32, vp])
register("oracle_cloud_shading_index_f64", None, [_cc, vp, i32, i32, vp])
register("oracle_cloud_shading_index_f32", None, [_cc, vp, i32, i32, vp])
register("oracle_cloud_reconstruct_f64", None, [_cc, vp, vp, vp, vp, vp, i32, i32, vp])
register("oracle_cloud_reconstruct_f32", None, [_cc, vp, vp, vp, vp, vp, i32, i32, vp])
register("oracle_cloud_upscale_f64", None, [_cc, vp, vp, vp, vp, i32, i32, vp])
register("oracle_cloud_upscale_f32", None, [_cc, vp, vp, vp, vp, i32, i32, vp])


def cloud_checkerboard_depth(depth):
    """[h/2, w/2] float32 from the [h, w] depth buffer."""
    d = np.ascontiguousarray(depth, dtype=np.float32)
    h, w = d.shape
    out = np.empty((h // 2, w // 2), np.float32)
    lib().oracle_cloud_checkerboard_depth(_ptr(d), w, h, _ptr(out))
    return out


def cloud_shading_index(common, checkerboard, precision="f64"):
    """[h/4, w/4, 2] (index, linear depth) from the [h/2, w/2] checkerboard."""
    cb = np.ascontiguousarray(checkerboard, dtype=np.float32)
    ch, cw = cb.shape
    out = np.empty((ch // 2, cw // 2, 2), _dt(precision))
    getattr(lib(), f"oracle_cloud_shading_index_{precision}")(C.byref(common), _ptr(cb), cw, ch, _ptr(out))
    return out


def cloud_reconstruct(common, checkerboard, index, render_bits, distance, previous_bits, w, h, precision="f64"):
    """[h/2, w/2, 4]: f64 values before the RGBA16F store, or (f32) the stored RGBA16F bits."""
    cb = np.ascontiguousarray(checkerboard, dtype=np.float32)
    ix = np.ascontiguousarray(index, dtype=np.float32)
    rb = np.ascontiguousarray(render_bits, dtype=np.uint16)
    dist = np.ascontiguousarray(distance, dtype=np.float32)
    prev = None if previous_bits is None else np.ascontiguousarray(previous_bits, dtype=np.uint16)
    out = np.empty((h // 2, w // 2, 4), np.float64 if precision == "f64" else np.uint16)
    getattr(lib(), f"oracle_cloud_reconstruct_{precision}")(C.byref(common), _ptr(cb), _ptr(ix), _ptr(rb), _ptr(dist),
                                                            _ptr(prev), w, h, _ptr(out))
    return out


def cloud_upscale(common, checkerboard, depth, reconstruct_bits, hdr_bits, precision="f64"):
    """[h, w, 4]: hdr composited under the upscaled reconstruction (f64 values, or f32 RGBA16F bits)."""
    cb = np.ascontiguousarray(checkerboard, dtype=np.float32)
    d = np.ascontiguousarray(depth, dtype=np.float32)
    rb = np.ascontiguousarray(reconstruct_bits, dtype=np.uint16)
    hb = np.ascontiguousarray(hdr_bits, dtype=np.uint16)
    h, w = d.shape
    out = np.empty((h, w, 4), np.float64 if precision == "f64" else np.uint16)
    getattr(lib(), f"oracle_cloud_upscale_{precision}")(C.byref(common), _ptr(cb), _ptr(d), _ptr(rb), _ptr(hb), w, h,
                                                        _ptr(out))
    return out


# The raymarch's dual-lobe phase for oracle raymarch calls.
register("oracle_bind_cloud_phase", None, [P(_abi.CloudPhase)])


def bind_cloud_phase(phase):
    """phase: (back_g, back_scale, forward_g), or None (the reference's lobes)."""
    if phase is None:
        lib().oracle_bind_cloud_phase(None)
    else:
        lib().oracle_bind_cloud_phase(C.byref(_abi.CloudPhase(*[float(x) for x in phase], 0.0)))


# The ray-scatter-visibility froxel bound for oracle raymarch / path-trace calls.
register("oracle_bind_shadow_froxel", None, [vp, i32, i32, i32])
_bound_froxel = None


def bind_shadow_froxel(froxel):
    """froxel: uint16 [d, h, w] R16 UNORM, or None (visibility 1)."""
    global _bound_froxel
    if froxel is None:
        _bound_froxel = None
        lib().oracle_bind_shadow_froxel(None, 0, 0, 0)
        return
    _bound_froxel = np.ascontiguousarray(froxel, dtype=np.uint16)
    d, h, w = _bound_froxel.shape
    lib().oracle_bind_shadow_froxel(_ptr(_bound_froxel), w, h, d)
