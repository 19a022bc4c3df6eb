"""Parameter presets and the synthetic workload definitions.

- `reference_earth()` is the `earth_.parameters` object of the reference's bin/config.json
  (numbers transcribed; they are the reference's parameter surface).
- `earth_lut_workload()` is the Earth LUT-set workload: the reference Earth with a 100 km
  atmosphere (top radius 6460 km), transmittance 256x64 / 40 steps, multiple scattering
  32x32 / 64 directions / 30 steps, sky view 192x108 / 30 steps, aerial perspective
  32x32x32 / 50.867 steps / 203 km, camera at 1 km altitude, 16 sun zenith angles over
  [0, 95] degrees stratified with numpy PCG64 seed 1.

  Mie absorption: the reference's 0.000444 / km (4.44e-7 / m).  The figure 4.4e-6 / m
  sometimes quoted for this preset is the Mie *extinction* (3.996e-6 + 0.444e-6 / m), which
  the reference also produces.  `mie_absorption_literal=True` gives the other reading.
"""
from __future__ import annotations

import numpy as np

from .atmosphere import AtmosphereParameters, AtmosphereRenderParameters, Camera


def reference_earth() -> AtmosphereParameters:
    """bin/config.json:62-107."""
    return AtmosphereParameters(
        solar_illuminance=(1.0, 1.0, 1.0),
        sun_angular_radius=0.26669999957084656,
        bottom_radius=6360.0,
        thickness=60.0,
        ground_albedo=(0.4000000059604645,) * 3,
        rayleigh_exponential_distribution=8.0,
        rayleigh_scattering_scale=0.03310000151395798,
        rayleigh_scattering=(0.17528699338436127, 0.4096069931983948, 1.0),
        mie_exponential_distribution=1.2000000476837158,
        mie_phase_g=0.800000011920929,
        mie_scattering_scale=0.003996000159531832,
        mie_scattering=(1.0, 1.0, 1.0),
        mie_absorption_scale=0.0004440000047907233,
        mie_absorption=(1.0, 1.0, 1.0),
        ozone_center_altitude=25.0,
        ozone_width=15.0,
        ozone_absorption_scale=0.0018810000037774444,
        ozone_absorption=(0.3455609977245331, 1.0, 0.04518900066614151),
        transmittance_steps=40.0,
        multiscattering_steps=30.0,
        multiscattering_mask=1.0,
    )


def earth_100km(mie_absorption_literal: bool = False) -> AtmosphereParameters:
    p = reference_earth()
    p.thickness = 100.0
    if mie_absorption_literal:
        p.mie_absorption_scale = 0.0044
    return p


def bench_camera(altitude_km: float = 1.0, pitch_deg: float = 10.0, aspect: float = 16.0 / 9.0) -> Camera:
    """World-space camera (y up, planet centre at (0, -R, 0)) looking along +x, pitched up.
    zNear / zFar as in bin/config.json:48-49."""
    return Camera(position=(0.0, altitude_km, 0.0), yaw=0.0, pitch=pitch_deg, aspect=aspect, fovy=45.0,
                  zNear=0.30000001192092896, zFar=50000.0)


def sun_zeniths(count: int = 16, max_deg: float = 95.0, seed: int = 1) -> np.ndarray:
    """Stratified sun zenith angles: max_deg * (k + u_k) / count, u_k ~ U[0, 1) (PCG64 seed)."""
    u = np.random.default_rng(seed).random(count)
    return (max_deg * (np.arange(count) + u) / count).astype(np.float32)


def earth_lut_workload():
    """Returns (atmosphere parameters, render parameters per sun angle, camera, sizes)."""
    atm = earth_100km()
    cam = bench_camera()
    vp = cam.view_projection()
    renders = []
    for theta in sun_zeniths():
        renders.append(AtmosphereRenderParameters(
            sun_direction_theta=float(theta), sun_direction_phi=180.0, star_luminance_scale=0.005,
            raymarching_steps=40.0, sky_view_lut_steps=30.0, aerial_perspective_lut_steps=50.867000579833984,
            aerial_perspective_lut_max_distance=203.0, camera_position=cam.position, view_projection=vp))
    sizes = {"transmittance": (256, 64), "multiscattering": (32, 32), "sky_view": (192, 108),
             "aerial_perspective": (32, 32, 32)}
    return atm, renders, cam, sizes


def integrator_steps_per_lut_set(sizes=None, t_steps=40, ms_steps=30, sv_steps=30, ap_steps=50.867000579833984):
    """Nominal integrator steps of one LUT set (T + MS + SV + AP), counting the float loop
    `for (i = 0.5; i < steps; ++i)` iterations per texel; AP froxel slices nearer than the
    camera (z = 0) march zero distance and are still counted as scheduled."""
    import math
    sizes = sizes or {"transmittance": (256, 64), "multiscattering": (32, 32), "sky_view": (192, 108),
                      "aerial_perspective": (32, 32, 32)}
    n = lambda s: math.ceil(s - 0.5)
    tw, th = sizes["transmittance"]
    mw, mh = sizes["multiscattering"]
    sw, sh = sizes["sky_view"]
    aw, ah, ad = sizes["aerial_perspective"]
    return (tw * th * n(t_steps) + mw * mh * 64 * n(ms_steps) + sw * sh * n(sv_steps) + aw * ah * ad * n(ap_steps))
