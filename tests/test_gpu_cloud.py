"""GPU parity of the cloud-texture generators and the volumetric-cloud raymarch.

Noise textures: UNORM8 outputs of the FP32 GPU generator against the FP64 oracle; a
value that lands within rounding distance of a UNORM8 step may round the other way, so
the bar is |diff| <= 1 LSB everywhere and < 0.5% of texels off by one.  Mip chains:
bit-exact (integer box filter).

Raymarch: the GPU and the oracle get the same texture, LUT and blue-noise bytes.  The
image metric (stated with every result): relative L2 error over the image
||gpu - oracle||_2 / ||oracle||_2 on RGB and on transmittance, bar 1e-3; plus the
fraction of pixels whose per-pixel relative error exceeds 1e-3 (floor 1e-3 * max),
bar 1% for the NEAREST-mip workload (a FP32/FP64 flip of a texel or mip choice moves a
whole sample) and 0.1% when every fetch is a continuous bilinear/trilinear one.  The
RGBA16F output's half rounding (<= 2^-11 relative) is inside these bars.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from skyrendering_amd import _abi, presets
from skyrendering_amd.cloud import CloudTextures, VolumetricCloudMinimalMaterial
from skyrendering_amd.scenes import CloudScene

pytestmark = pytest.mark.gpu


def _lsb_check(gpu, ref, what):
    diff = np.abs(gpu.astype(np.int32) - ref.astype(np.int32))
    frac = float(np.mean(diff > 0))
    print(f"{what}: max |diff| {diff.max()} LSB, fraction off-by-one {frac:.2e} of {diff.size}")
    assert diff.max() <= 1 and frac < 5e-3


def test_noise_textures_match_oracle(ctx):
    m = presets.cloud_material_reference()
    tex = CloudTextures(ctx, 512, 128, 128).generate(m)
    cm, det, disp = tex.download()
    n = 512 * 512
    ref = oracle.cloud_map_texels(m.cloud_map_density.to_c(), m.cloud_map_height.to_c(), 512, np.arange(n))
    _lsb_check(cm[:2 * n].reshape(n, 2), ref, "cloud map 512^2 RG8")
    sample = np.random.default_rng(0).choice(128 ** 3, 20000, replace=False)
    ref = oracle.detail_texels(m.detail_perlin.to_c(), m.detail_worley.to_c(), 128, sample)
    _lsb_check(det[sample], ref, "detail 128^3 R8 (20000-texel sample)")
    n = 128 * 128
    ref = oracle.displacement_texels(m.displacement_perlin.to_c(), 128, np.arange(n))
    _lsb_check(disp[:4 * n].reshape(n, 4), ref, "displacement 128^2 RGBA8")


def test_mip_chains_bit_exact(ctx):
    m = presets.cloud_material_reference()
    tex = CloudTextures(ctx, 512, 128, 128).generate(m)
    for name, chain, dims, ch, size, levels in [("cloud map", tex.cloud_map, 2, 2, 512, tex.cloud_map_levels),
                                                ("detail", tex.detail, 3, 1, 128, tex.detail_levels),
                                                ("displacement", tex.displacement, 2, 4, 128, tex.displacement_levels)]:
        got = chain.download(np.uint8, (chain.nbytes,))
        base = size * size * (size if dims == 3 else 1) * ch
        ref = got.copy()
        ref[base:] = 0
        ref = oracle.generate_mipmaps(ref, dims, ch, size, size, size, levels)
        assert np.array_equal(got, ref), name


def _oracle_inputs(scene: CloudScene):
    a, r = scene.atmosphere, scene.renderer
    apl, apt = r.aerial_perspective()
    kw = dict(T=a.transmittance(), ap_l=apl, ap_t=apt, blue_noise=scene.blue_noise_host,
              image_size=(scene.width, scene.height))
    if scene.textures is not None:
        cm, det, disp = scene.textures.download()
        kw.update(cloud_map=cm, detail=det, displacement=disp)
    return oracle.HostCloudInputs(scene.mat_block, **kw)


def _compare(scene, pixel_tol_frac, label, sample=None):
    scene.run(count_steps=True)
    rgba, dist = scene.image()
    gpu_steps = scene.executed_steps()
    w, h = scene.width, scene.height
    px = np.arange(w * h) if sample is None else sample
    ref, ref_dist, ref_steps = oracle.cloud_raymarch(scene.atmosphere.buffer, scene.common, scene.cloud_block,
                                                     _oracle_inputs(scene), w, px)
    g = rgba.reshape(-1, 4)[px].astype(np.float64)
    l2_rgb = np.linalg.norm(g[:, :3] - ref[:, :3]) / np.linalg.norm(ref[:, :3])
    l2_t = np.linalg.norm(g[:, 3] - ref[:, 3]) / np.linalg.norm(ref[:, 3])
    floor = 1e-3 * np.abs(ref[:, :4]).max(axis=0)
    per = np.max(np.abs(g - ref) / np.maximum(np.abs(ref), floor), axis=1)
    frac = float(np.mean(per > 1e-3))
    d_err = np.abs(dist.reshape(-1)[px] - ref_dist) / np.maximum(ref_dist, 1e-3)
    print(f"{label}: rel L2 rgb {l2_rgb:.2e}, T {l2_t:.2e}; pixels > 1e-3: {frac:.2e} (max {per.max():.2e}); "
          f"distance median rel {np.median(d_err):.2e}; steps gpu {gpu_steps} oracle {int(ref_steps.sum())}")
    assert l2_rgb < 1e-3 and l2_t < 1e-3
    assert frac < pixel_tol_frac
    if sample is None:
        assert abs(gpu_steps - int(ref_steps.sum())) <= 1e-3 * ref_steps.sum()
    return gpu_steps


def test_raymarch_benchmark_workload_parity(ctx):
    """The 1080p workload's scene (weather map, detail volume, LOD rule, 64 steps, 6 light
    steps) at 480x270, every pixel."""
    scene = CloudScene(ctx, 480, 270)
    _compare(scene, 1e-2, "Material1 480x270 (NEAREST mips)")


def test_raymarch_1080p_sample_parity(ctx):
    scene = CloudScene(ctx, 1920, 1080)
    sample = np.random.default_rng(5).choice(1920 * 1080, 20000, replace=False)
    _compare(scene, 1e-2, "Material1 1920x1080 (20000-pixel sample)", sample=sample)


def test_raymarch_continuous_fetches_parity(ctx):
    """LOD bias -30 keeps every fetch on level 0 (bilinear / trilinear), so the only
    discontinuities left are the early exit and the sigma < 1e-5 skip."""
    scene = CloudScene(ctx, 320, 180, lod_bias=-30.0)
    _compare(scene, 1e-3, "Material1 320x180 magnification only")


def test_raymarch_noise_cloud_map_and_reference_blue_noise(ctx, blue_noise_64):
    """The reference's own texture pipeline: the NoiseGen cloud map instead of the uniform
    weather map, and the 64x64 blue-noise tile of data/BlueNoise (pos & 0x3f)."""
    scene = CloudScene(ctx, 320, 180, weather="noise", blue_noise=blue_noise_64, frame_id=7)
    _compare(scene, 1e-2, "Material1 NoiseGen cloud map, blue-noise fixture, frame 7")


def test_raymarch_minimal_material_parity(ctx):
    scene = CloudScene(ctx, 320, 180, material=VolumetricCloudMinimalMaterial(density=3.0))
    _compare(scene, 1e-3, "Minimal material")


def test_raymarch_rejects_bad_input(ctx):
    scene = CloudScene(ctx, 64, 36)
    bad = _abi.CloudMaterial()
    C.memmove(C.byref(bad), C.byref(scene.mat_block), C.sizeof(bad))
    bad.cloud_map_size = 500  # not a power of two
    a = scene.atmosphere
    L = _abi.lib()
    rc = L.skyhip_cloud_raymarch(ctx.handle, C.byref(a.buffer), C.byref(scene.common), C.byref(scene.cloud_block),
                                 C.byref(bad), a.transmittance_texture.ptr, 256, 64,
                                 scene.renderer.aerial_perspective_luminance_texture.ptr,
                                 scene.renderer.aerial_perspective_transmittance_texture.ptr, 32, 32, 32,
                                 scene.blue_noise.ptr, 128, None, None, scene.out_rgba.ptr, scene.out_dist.ptr, 64, 36,
                                 None)
    assert rc == _abi.ERR_INVALID_SIZE
    bad.kind = 9
    rc = L.skyhip_cloud_raymarch(ctx.handle, C.byref(a.buffer), C.byref(scene.common), C.byref(scene.cloud_block),
                                 C.byref(bad), a.transmittance_texture.ptr, 256, 64,
                                 scene.renderer.aerial_perspective_luminance_texture.ptr,
                                 scene.renderer.aerial_perspective_transmittance_texture.ptr, 32, 32, 32,
                                 scene.blue_noise.ptr, 128, None, None, scene.out_rgba.ptr, scene.out_dist.ptr, 64, 36,
                                 None)
    assert rc == _abi.ERR_INVALID_ARGUMENT
