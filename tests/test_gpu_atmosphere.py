"""GPU parity of the atmosphere LUT kernels against the FP64 oracle.

Tolerance: max per-element relative error <= 1e-3 (|gpu - oracle| / max(|oracle|,
1e-6 * max|oracle|)), RGB channels.  Each pass is fed the same upstream LUT bytes as the
oracle, so errors do not compound across passes; a chained test covers the whole set.
"""
import numpy as np
import pytest

import oracle
from skyrendering_amd import Atmosphere, AtmosphereRenderer, AtmosphereRenderInitParameters, presets
from skyrendering_amd.atmosphere import AtmosphereParameters, AtmosphereRenderParameters

from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-3


def _render(theta, phi=180.0, altitude=1.0, sv_steps=30.0):
    cam = presets.bench_camera(altitude_km=altitude)
    return AtmosphereRenderParameters(sun_direction_theta=theta, sun_direction_phi=phi, sky_view_lut_steps=sv_steps,
                                      aerial_perspective_lut_steps=50.867000579833984,
                                      aerial_perspective_lut_max_distance=203.0, camera_position=cam.position,
                                      view_projection=cam.view_projection())


@pytest.mark.parametrize("preset,size", [("earth_100km", (256, 64)), ("reference_earth", (256, 64)),
                                         ("earth_100km", (128, 32))])
def test_transmittance_lut_parity(ctx, preset, size):
    p = getattr(presets, preset)()
    atm = Atmosphere(ctx, transmittance_size=size)
    atm.update_luts(p)
    T = atm.transmittance()
    ref = oracle.transmittance_lut(atm.buffer, *size)
    e, rms, at = rel_err(T[..., :3], ref[..., :3])
    print(f"T {preset} {size}: max rel {e:.3e} rms {rms:.3e} at {at}")
    assert e < TOL
    assert np.all(T[..., 3] == 1.0)


def test_multiscattering_lut_parity(ctx):
    p = presets.earth_100km()
    atm = Atmosphere(ctx)
    atm.update_luts(p)
    T = atm.transmittance()
    MS = atm.multiscattering()
    ref = oracle.multiscattering_lut(atm.buffer, T, 32, 32)
    e, rms, at = rel_err(MS[..., :3], ref[..., :3])
    print(f"MS: max rel {e:.3e} rms {rms:.3e} at {at}")
    assert e < TOL


@pytest.mark.parametrize("theta", [0.0, 45.0, 85.0, 95.0])
def test_sky_view_lut_parity(ctx, theta):
    p = presets.earth_100km()
    atm = Atmosphere(ctx)
    atm.update_luts(p)
    rend = AtmosphereRenderer(ctx, AtmosphereRenderInitParameters(sky_view_lut_size=(192, 108)))
    rend.render(atm, p, _render(theta), aerial_perspective=False)
    L, Tr = rend.sky_view()
    refL, refT = oracle.sky_view_lut(atm.buffer, rend.buffer, atm.transmittance(), atm.multiscattering(), 192, 108)
    eL, rmsL, atL = rel_err(L[..., :3], refL[..., :3])
    eT, rmsT, atT = rel_err(Tr[..., :3], refT[..., :3])
    print(f"SV theta={theta}: L max rel {eL:.3e} rms {rmsL:.3e} at {atL}; T max rel {eT:.3e} rms {rmsT:.3e}")
    assert eL < TOL and eT < TOL


def test_sky_view_from_space_and_dither(ctx, blue_noise_64):
    p = presets.earth_100km()
    atm = Atmosphere(ctx)
    atm.update_luts(p)
    bn = ctx.upload(blue_noise_64)
    rend = AtmosphereRenderer(ctx, AtmosphereRenderInitParameters(sky_view_lut_size=(96, 54),
                                                                  sky_view_lut_dither_sample_point_enable=True),
                              blue_noise=bn, blue_noise_size=64)
    for altitude in [1.0, 250.0]:
        rend.render(atm, p, _render(30.0, altitude=altitude, sv_steps=40.0), aerial_perspective=False)
        L, Tr = rend.sky_view()
        refL, refT = oracle.sky_view_lut(atm.buffer, rend.buffer, atm.transmittance(), atm.multiscattering(), 96, 54,
                                         blue_noise=blue_noise_64)
        eL, _, atL = rel_err(L[..., :3], refL[..., :3])
        eT, _, _ = rel_err(Tr[..., :3], refT[..., :3])
        print(f"SV altitude {altitude}: L {eL:.3e} at {atL}, T {eT:.3e}")
        assert eL < TOL and eT < TOL


@pytest.mark.parametrize("theta", [20.0, 88.0])
def test_aerial_perspective_lut_parity(ctx, theta):
    p = presets.earth_100km()
    atm = Atmosphere(ctx)
    atm.update_luts(p)
    rend = AtmosphereRenderer(ctx, AtmosphereRenderInitParameters())
    rend.render(atm, p, _render(theta), sky_view=False)
    L, Tr = rend.aerial_perspective()
    refL, refT = oracle.aerial_perspective_lut(atm.buffer, rend.buffer, atm.transmittance(), atm.multiscattering(),
                                               32, 32, 32)
    eL, rmsL, atL = rel_err(L[..., :3], refL[..., :3])
    eT, rmsT, atT = rel_err(Tr[..., :3], refT[..., :3])
    print(f"AP theta={theta}: L max rel {eL:.3e} rms {rmsL:.3e} at {atL}; T max rel {eT:.3e}")
    assert eL < TOL and eT < TOL
    # Slice 0 marches zero distance: L = 0, T = 1 exactly.
    assert np.all(L[0, ..., :3] == 0) and np.all(Tr[0, ..., :3] == 1)


def test_lut_chain_end_to_end_against_oracle_chain(ctx):
    """GPU T -> MS -> SV against the oracle's own FP64 chain (LUTs stored as FP32, as the
    reference's RGBA32F textures are)."""
    p = presets.earth_100km()
    atm = Atmosphere(ctx)
    atm.update_luts(p)
    rend = AtmosphereRenderer(ctx, AtmosphereRenderInitParameters(sky_view_lut_size=(192, 108)))
    rend.render(atm, p, _render(60.0), aerial_perspective=False)
    L, _ = rend.sky_view()
    T64 = oracle.transmittance_lut(atm.buffer, 256, 64).astype(np.float32)
    MS64 = oracle.multiscattering_lut(atm.buffer, T64, 32, 32).astype(np.float32)
    refL, _ = oracle.sky_view_lut(atm.buffer, rend.buffer, T64, MS64, 192, 108)
    e, rms, at = rel_err(L[..., :3], refL[..., :3])
    print(f"chain SV: max rel {e:.3e} rms {rms:.3e} at {at}")
    assert e < TOL


def test_update_luts_rejects_bad_sizes(ctx):
    import ctypes as C
    from skyrendering_amd import _abi
    p = presets.earth_100km()
    buf = ctx.alloc(16 * 4)
    rc = _abi.lib().skyhip_atmosphere_update_luts(ctx.handle, C.byref(p.to_c()), buf.ptr, 1, 64, buf.ptr, 32, 32)
    assert rc == _abi.ERR_INVALID_SIZE


def test_lut_set_single_call_matches_oracle(ctx):
    """skyhip_atmosphere_lut_set (T -> MS -> SV + AP in one launch, marches split over lanes) at the Earth
    LUT-set sizes."""
    from skyrendering_amd.lutset import LutSet
    atm_p, renders, cam, sizes = presets.earth_lut_workload()
    ls = LutSet(ctx, sizes)
    for r in (renders[0], renders[-1]):
        ls.build(atm_p, r)
        out = ls.download()
        refT = oracle.transmittance_lut(ls.atm_block, 256, 64)
        refMS = oracle.multiscattering_lut(ls.atm_block, out["T"], 32, 32)
        refL, refTr = oracle.sky_view_lut(ls.atm_block, ls.render_block, out["T"], out["MS"], 192, 108)
        apL, apT = oracle.aerial_perspective_lut(ls.atm_block, ls.render_block, out["T"], out["MS"], 32, 32, 32)
        errs = {k: rel_err(a[..., :3], b[..., :3])[0] for k, a, b in [
            ("T", out["T"], refT), ("MS", out["MS"], refMS), ("SV_L", out["SV_L"], refL),
            ("SV_T", out["SV_T"], refTr), ("AP_L", out["AP_L"], apL), ("AP_T", out["AP_T"], apT)]}
        print(f"LUT set theta={r.sun_direction_theta:.2f}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        assert max(errs.values()) < TOL
