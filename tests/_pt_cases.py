"""Small path-tracing scenes shared by the CPU oracle tests and the GPU parity tests."""
import math

import numpy as np

from skyrendering_amd import _abi, presets
from skyrendering_amd.atmosphere import AtmosphereRenderParameters
from skyrendering_amd.pathtrace import PathTracingInitParam


def straight_up_common(sun_zenith_deg=40.0, hb=1.5, ht=4.0, R=6360.0):
    """A cloud common block whose every pixel looks straight up (+z) from the local origin:
    uInvMVP sends every NDC point to (0, 0, 1000)."""
    c = _abi.CloudCommonBufferData()
    for i in range(16):
        c.uInvMVP[i] = 0.0
    c.uInvMVP[14] = 1000.0
    c.uInvMVP[15] = 1.0
    for i in range(3):
        c.uCameraPos[i] = 0.0
    z = math.radians(sun_zenith_deg)
    sun = np.array([math.sin(z), 0.0, math.cos(z)], np.float32)
    for i in range(3):
        c.uSunDirection[i] = float(sun[i])
    c.uBottomAltitude, c.uTopAltitude = hb, ht
    c.uEarthRadius = R
    c.uAerialPerspectiveLutMaxDistance = 203.0
    return c


def render_block(theta=40.0):
    cam = presets.bench_camera(aspect=1.0)
    p = AtmosphereRenderParameters(sun_direction_theta=theta, sun_direction_phi=180.0, sky_view_lut_steps=30.0,
                                   aerial_perspective_lut_steps=50.867000579833984,
                                   aerial_perspective_lut_max_distance=203.0, camera_position=cam.position,
                                   view_projection=cam.view_projection())
    return p.buffer_data(6360.0)


def unit_luts(sv_value=1.0):
    """T = 1, sky view = sv_value (luminance) / 1 (transmittance), AP luminance 0 / transmittance 1."""
    T = np.ones((4, 4, 4), np.float32)
    sv_l = np.full((8, 8, 4), sv_value, np.float32)
    sv_t = np.ones((8, 8, 4), np.float32)
    ap_l = np.zeros((2, 2, 2, 4), np.float32)
    ap_t = np.ones((2, 2, 2, 4), np.float32)
    return T, sv_l, sv_t, ap_l, ap_t


def minimal_material(sigma):
    m = _abi.CloudMaterial()
    m.kind = _abi.MATERIAL_MINIMAL
    m.minimal.uDensity = sigma
    return m


def params(sigma_t_max, max_bounces=1, env="OFF", albedo=1.0, importance=True, ratio=0.7, half_width=100.0,
           prng="PCGHash", model=None, gf=0.85, gb=-0.15):
    init = PathTracingInitParam(max_bounces=max_bounces, region_box_half_width=half_width,
                                importance_sampling=importance, forward_phase_g=gf, back_phase_g=gb,
                                forward_scattering_ratio=ratio, prng=prng, environment_lighting=env, albedo=albedo)
    model = np.eye(4, dtype=np.float32).reshape(-1) if model is None else model
    return init.to_c(sigma_t_max, model)


def hg(c, g):
    return (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * c) ** 1.5)


F32 = lambda x: float(np.float32(x))  # the parameter blocks hold FP32


def hg_mixture(c, gf=F32(0.85), gb=F32(-0.15), ratio=F32(0.7)):
    return ratio * hg(c, gf) + (1 - ratio) * hg(c, gb)


def hg_cdf(mu, g):
    """P(cos theta <= mu) of Henyey-Greenstein (integral of 2 pi hg over [-1, mu])."""
    return (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * mu) - 1 / (1 + g))


def small_grid_workload(dims=(64, 64, 64), density=2.0, spp=16):
    w = presets.path_trace_workload()
    w["grid_dims"] = dims
    w["density"] = density
    w["spp"] = spp
    return w


def cpu_scene(workload, width, height, bricks_note=False):
    """The path-tracing workload's inputs built on the CPU with the oracle (FP32 LUT storage):
    returns (atmosphere block, render block, common block, params, HostPathTraceInputs)."""
    import oracle
    w = workload
    cam = w["camera"]
    cam.aspect = width / height
    atm = w["atmosphere"].buffer_data()
    rp = AtmosphereRenderParameters(sun_direction_theta=w["sun_theta"], sun_direction_phi=w["sun_phi"],
                                    sky_view_lut_steps=30.0, aerial_perspective_lut_steps=50.867000579833984,
                                    aerial_perspective_lut_max_distance=203.0, camera_position=cam.position,
                                    view_projection=cam.view_projection())
    render = rp.buffer_data(w["atmosphere"].bottom_radius)
    T = oracle.transmittance_lut(atm, 256, 64).astype(np.float32)
    MS = oracle.multiscattering_lut(atm, T, 32, 32).astype(np.float32)
    sv_l, sv_t = oracle.sky_view_lut(atm, render, T, MS, 192, 108)
    ap_l, ap_t = oracle.aerial_perspective_lut(atm, render, T, MS, 32, 32, 32)
    sun = np.array(render.sun_direction[:], np.float32)
    common, _, model = w["cloud"].update(cam, w["atmosphere"].bottom_radius, sun, 203.0, 0)
    grid = oracle.voxel_grid_f16(w["grid"].to_c(), w["grid_dims"])
    center = presets.grid_center_ahead(model, cam, w["grid_distance_km"])
    ext = w["grid_extent_km"]
    m = _abi.CloudMaterial()
    m.kind = _abi.MATERIAL_VOXEL_F16
    m.voxel.uSampleFrequency[0] = m.voxel.uSampleFrequency[1] = 1.0 / ext
    m.voxel.uSampleBias[0], m.voxel.uSampleBias[1] = 0.5 - center[0] / ext, 0.5 - center[1] / ext
    m.voxel.uDensity = w["density"]
    for k in range(3):
        m.voxel_grid_dim[k] = w["grid_dims"][k]
    gmax = float(grid.view(np.float16).astype(np.float32).max())
    prm = w["init"].to_c(gmax * w["density"], model)
    inputs = oracle.HostPathTraceInputs(m, T, sv_l.astype(np.float32), sv_t.astype(np.float32),
                                        ap_l.astype(np.float32), ap_t.astype(np.float32), grid=grid)
    return atm, render, common, prm, inputs
