"""The oracle's environment-luminance cube and IBL precomputation, pinned by geometry and
closed forms (the reference has no tests of its own).

- Face tables: every texel's direction (ConvertCubUvToDir with the passes' v flip) selects
  the same face and texel through the OpenGL cube-map face table; face centres point along
  the axes.
- SH9 (EnvRadianceSH.comp): a constant sky c gives L00 = c Y00 4 pi and small higher
  terms; the sky 1 + k.w/2 gives the degree-1 coefficients Y1 k/2 4 pi / 3 on the right
  axes (index 1: y, 2: z, 3: x), which a transposed index would break.
- Prefilter (PrefilterRadiance.comp): roughness 0 reproduces the source texels exactly; a
  constant sky stays constant at every roughness; for the linear sky the result stays an
  affine function of k.R (the GGX lobe is symmetric about R).
- Environment cube: with a constant sky-view luminance and no ground in view the texels
  equal it; the mip chain is the 2x2 box of the level below.
"""
import math

import numpy as np

import oracle
from skyrendering_amd import presets
from skyrendering_amd.atmosphere import AtmosphereRenderParameters

Y00, Y1N = float(np.float32(0.282095)), float(np.float32(0.488603))  # the shader's float constants


def _cube_from(fn, w):
    """A full cube (uint16 bits) whose level-0 texel values are fn(direction), mips by box."""
    lvl0 = np.zeros((6, w, w, 4), np.float16)
    for f in range(6):
        for y in range(w):
            for x in range(w):
                lvl0[f, y, x, :3] = fn(oracle.cube_texel_dir(f, x, y, w))
    flat = np.zeros((oracle.cube_texels(w, oracle.cube_levels(w)), 4), np.uint16)
    flat[:6 * w * w] = lvl0.reshape(-1, 4).view(np.uint16)
    return oracle.cube_mips(flat, w)


def test_cube_face_tables_round_trip():
    for w in (8, 128):
        for f in range(6):
            for (x, y) in ((0, 0), (w - 1, 0), (3, w - 2), (w // 2, w // 3), (w - 1, w - 1)):
                d = oracle.cube_texel_dir(f, x, y, w)
                face, s, t = oracle.cube_face_st(d)
                assert face == f
                assert abs(s * w - 0.5 - x) < 1e-9 and abs(t * w - 0.5 - y) < 1e-9
    axes = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for f, a in enumerate(axes):  # the centre of an odd-sized face is its axis
        assert np.allclose(oracle.cube_texel_dir(f, 1, 1, 3), a, atol=1e-12)


def test_sh9_closed_forms():
    w = 32
    c = 0.75
    sh = oracle.env_radiance_sh(_cube_from(lambda d: (c, c, c), w), w)
    assert abs(sh[0, 0] - c * Y00 * 4 * math.pi) < 1e-9
    assert np.abs(sh[1:, :3]).max() < 1e-2 * c and np.all(sh[:, 3] == 0)
    k = np.array([0.3, -0.5, 0.8])
    sh = oracle.env_radiance_sh(_cube_from(lambda d: (1 + 0.5 * np.dot(k, d),) * 3, w), w)
    want = {1: k[1], 2: k[2], 3: k[0]}
    for idx, kv in want.items():
        expect = Y1N * 0.5 * kv * 4 * math.pi / 3
        assert abs(sh[idx, 0] - expect) < 2e-3 * abs(Y1N * 0.5 * 4 * math.pi / 3), (idx, sh[idx, 0], expect)
    assert abs(sh[0, 0] - Y00 * 4 * math.pi) < 1e-3
    assert np.abs(sh[4:, 0]).max() < 1e-2


def test_prefilter_closed_forms():
    w = 16
    k = np.array([0.3, -0.5, 0.8])
    env = _cube_from(lambda d: (1 + 0.5 * np.dot(k, d),) * 3, w)
    pre = oracle.prefiltered_radiance(env, w, w, 4)
    assert np.array_equal(pre[:6 * w * w], env[:6 * w * w])  # roughness 0 is the source texel
    flat = _cube_from(lambda d: (0.5, 0.5, 0.5), w)
    pc = oracle.prefiltered_radiance(flat, w, w, 4).view(np.float16).astype(np.float64)
    assert np.all(pc[:, :3] == 0.5)
    # Linear sky: each level is affine in k.R; the slope shrinks as the lobe widens.
    levels = [0, 6 * w * w, 6 * w * w + 6 * (w // 2) ** 2, 6 * w * w + 6 * (w // 2) ** 2 + 6 * (w // 4) ** 2]
    vals = pre.view(np.float16).astype(np.float64)
    slopes = []
    for lvl in range(1, 4):
        wl = w >> lvl
        kr = np.array([np.dot(k, oracle.cube_texel_dir(f, x, y, wl)) for f in range(6) for y in range(wl)
                       for x in range(wl)])
        v = vals[levels[lvl]:levels[lvl] + 6 * wl * wl, 0]
        A = np.stack([np.ones_like(kr), kr], 1)
        coef, res, *_ = np.linalg.lstsq(A, v, rcond=None)
        resid = np.abs(A @ coef - v).max()
        slopes.append(coef[1])
        assert resid < 2e-2, (lvl, resid)
    assert 0.5 >= slopes[0] > slopes[1] > slopes[2] > 0.0, slopes


def test_environment_cube_constant_sky():
    w = 16
    p = presets.earth_100km()
    atm = p.buffer_data()
    cam = presets.bench_camera()
    rp = AtmosphereRenderParameters(sun_direction_theta=45.0, camera_position=cam.position,
                                    view_projection=cam.view_projection())
    render = rp.buffer_data(p.bottom_radius)
    T = np.ones((16, 64, 4), np.float32)
    sv_l = np.full((18, 32, 4), 0.25, np.float32)
    sv_t = np.ones((18, 32, 4), np.float32)
    cube = oracle.environment_cube(atm, render, T, sv_l, sv_t, w)
    lv0 = cube[:6 * w * w].view(np.float16).astype(np.float64).reshape(6, w, w, 4)
    up = np.array(render.up_direction[:])
    above = np.array([[[np.dot(up, oracle.cube_texel_dir(f, x, y, w)) > 0.05 for x in range(w)] for y in range(w)]
                      for f in range(6)])
    assert above.sum() > 6 * w * w * 0.4
    assert np.all(lv0[above][:, :3] == 0.25) and np.all(lv0[..., 3] == 0)
    assert np.all(lv0[~above][:, :3] >= 0.25)  # ground adds to the sky below the horizon
    assert np.array_equal(oracle.cube_mips(cube, w), cube)
