"""GPU parity of the cloud's checkerboard depth, shading index, temporal reconstruction and
upscale (skyhip_cloud_checkerboard_depth / _shading_index / _reconstruct / _upscale).

These passes are selections and short FP32 expressions; the kernels evaluate them in the
shaders' order without FMA contraction, as the oracle's FP32 path does (x86-64-v2, no FMA),
so the bar is bit-exact: identical floats for the checkerboard and the (index, linear depth)
image, identical RGBA16F bits for the reconstruction and the composite.  The oracle's FP64
path is checked beside it (relative error of the reconstruction where both take the same
branches).  Inputs: a synthetic depth buffer (sky, a sloping terrain with ripples, near boxes),
random quarter-resolution samples, distances and history, and a camera that moved between
frames (sub-texel reprojection); the tile rows of the fused reconstruction + upscale end in a
partial tile at this height.  The chain test runs the real quarter-resolution raymarch in index
mode (VolumetricCloud::Render) for two frames, fused and one launch per pass, and checks each
pass against the oracle on the GPU's own inputs.
"""
import numpy as np
import pytest

import oracle
from skyrendering_amd import presets
from skyrendering_amd.cloud import CloudReconstruction, VolumetricCloud, cloud_reproject_matrix
from skyrendering_amd.scenes import CloudScene

pytestmark = pytest.mark.gpu

W, H = 320, 184


def _depth_buffer(common, w, h, seed=0):
    return presets.synthetic_depth_buffer(common.uLinearDepthParam, w, h, seed)


def _frames():
    cloud = VolumetricCloud()
    sun = np.array([0.3, 0.8, 0.2], np.float32)
    sun /= np.linalg.norm(sun)
    cam1 = presets.bench_camera(aspect=W / H)
    c1 = cloud.update(cam1, 6360.0, sun, 203.0, 0)[0]
    mvp1 = cloud_reproject_matrix(cam1, 6360.0, c1)
    cam2 = presets.bench_camera(aspect=W / H)
    cam2.position = (0.04, 1.01, -0.02)
    cam2.pitch = 10.3
    c2 = cloud.update(cam2, 6360.0, sun, 203.0, 1)[0]
    cloud_reproject_matrix(cam2, 6360.0, c2, previous_mvp=mvp1, previous_camera_position=cam1.position,
                           additional_delta=(0.003, -0.002))
    return c1, c2


def _half_bits(a):
    return np.asarray(a, np.float32).astype(np.float16).view(np.uint16)


def _same_bits(gpu, ref, what):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    diff = int(np.count_nonzero(gpu != ref))
    print(f"{what}: {gpu.size - diff}/{gpu.size} identical")
    if diff:
        i = np.argwhere(gpu != ref)[:5]
        print("  first differences:", [(tuple(j), gpu[tuple(j)], ref[tuple(j)]) for j in i])
    assert diff == 0, what


def test_passes_bit_exact_against_oracle(ctx):
    c1, c2 = _frames()
    depth = _depth_buffer(c2, W, H)
    rng = np.random.default_rng(4)
    rec = CloudReconstruction(ctx, (W, H))
    d_depth = ctx.upload(depth)
    rec.checkerboard_depth(d_depth)
    for base in (0, 3):
        c2.uBaseShadingIndex = base
        rec.shading_index(c2)
        ctx.synchronize()
        cb = rec.checkerboard.download(np.float32, (H // 2, W // 2))
        _same_bits(cb, oracle.cloud_checkerboard_depth(depth), "checkerboard depth")
        idx = rec.index.download(np.float32, (H // 4, W // 4, 2))
        _same_bits(idx, oracle.cloud_shading_index(c2, cb, "f32"), f"shading index (base {base})")
        ref64 = oracle.cloud_shading_index(c2, cb, "f64")
        assert np.mean(ref64[..., 0] == idx[..., 0]) > 0.999
    q = (H // 4, W // 4)
    render = _half_bits(np.concatenate([rng.random(q + (3,)) * 5.0, rng.random(q + (1,))], -1))
    dist = (1.0 + 29.0 * rng.random(q)).astype(np.float32)
    prev = _half_bits(np.concatenate([rng.random((H // 2, W // 2, 3)) * 5.0, rng.random((H // 2, W // 2, 1))], -1))
    rec.render_texture.upload(render)
    rec.distance.upload(dist)
    rec.reconstruct[1].upload(prev)
    for frames, history in ((0, None), (1, prev)):
        rec.frames = frames
        rec.reconstruct_pass(c2)
        ctx.synchronize()
        got = rec.reconstruct[0].download(np.uint16, (H // 2, W // 2, 4))
        ref = oracle.cloud_reconstruct(c2, cb, idx, render, dist, history, W, H, "f32")
        _same_bits(got, ref, f"reconstruction ({'with' if history is not None else 'no'} history)")
        r64 = oracle.cloud_reconstruct(c2, cb, idx, render, dist, history, W, H, "f64")
        g = got.view(np.float16).astype(np.float64)
        rel = np.abs(g - r64) / np.maximum(np.abs(r64), 1e-3)
        print(f"  vs FP64: median rel {np.median(rel):.2e}, fraction > 1e-3 {np.mean(rel > 1e-3):.2e}")
        assert np.mean(rel > 1e-3) < 1e-3
    hdr = _half_bits(np.concatenate([rng.random((H, W, 3)) * 20.0, np.ones((H, W, 1))], -1))
    d_hdr = ctx.upload(hdr)
    rec.upscale(c2, d_depth, d_hdr)
    ctx.synchronize()
    _same_bits(d_hdr.download(np.uint16, (H, W, 4)), oracle.cloud_upscale(c2, cb, depth, got, hdr, "f32"),
               "upscale composite")
    # The fused launches (checkerboard + index; reconstruction + upscale through LDS tiles, with
    # a partial tile row at this height): the same bits.
    fused = CloudReconstruction(ctx, (W, H))
    fused.checkerboard_index(c2, d_depth)
    ctx.synchronize()
    _same_bits(fused.checkerboard.download(np.float32, (H // 2, W // 2)), cb, "fused checkerboard")
    _same_bits(fused.index.download(np.float32, (H // 4, W // 4, 2)), oracle.cloud_shading_index(c2, cb, "f32"),
               "fused shading index (base 3)")
    fused.index.upload(idx)
    fused.render_texture.upload(render)
    fused.distance.upload(dist)
    fused.reconstruct[1].upload(prev)
    for frames, history in ((0, None), (1, prev)):
        fused.frames = frames
        d_hdr.upload(hdr)
        fused.reconstruct_upscale(c2, d_depth, d_hdr)
        ctx.synchronize()
        got = fused.reconstruct[0].download(np.uint16, (H // 2, W // 2, 4))
        _same_bits(got, oracle.cloud_reconstruct(c2, cb, idx, render, dist, history, W, H, "f32"),
                   f"fused reconstruction ({'with' if history is not None else 'no'} history)")
        _same_bits(d_hdr.download(np.uint16, (H, W, 4)), oracle.cloud_upscale(c2, cb, depth, got, hdr, "f32"),
                   "fused upscale composite")


@pytest.mark.parametrize("fused", [True, False])
def test_render_chain_two_frames(ctx, fused):
    scene = CloudScene(ctx, W, H)
    c = scene.common
    mvp = cloud_reproject_matrix(scene.camera, scene.earth.bottom_radius, c)
    depth = _depth_buffer(c, W, H, seed=1)
    d_depth = ctx.upload(depth)
    rec = CloudReconstruction(ctx, (W, H))
    a, r = scene.atmosphere, scene.renderer
    aw, ah = r.init.aerial_perspective_lut_size
    hdr0 = _half_bits(np.concatenate([np.full((H, W, 3), 2.0), np.ones((H, W, 1))], -1))
    d_hdr = ctx.upload(hdr0)
    for frame in range(2):
        c.uBaseShadingIndex = frame & 3
        c.uFrameID = float(frame)
        cloud_reproject_matrix(scene.camera, scene.earth.bottom_radius, c, previous_mvp=mvp,
                               previous_camera_position=scene.camera.position)
        d_hdr.upload(hdr0)
        history = rec.reconstruct[1].download(np.uint16, (H // 2, W // 2, 4)) if frame else None
        rec.render(a.buffer, c, scene.cloud_block, scene.mat_block, a.transmittance_texture, a.transmittance_size,
                   r.aerial_perspective_luminance_texture, r.aerial_perspective_transmittance_texture,
                   (aw, ah, r.init.aerial_perspective_lut_depth), scene.blue_noise, scene.blue_noise_host.shape[0],
                   d_depth, d_hdr, fused=fused)
        ctx.synchronize()
        cb = rec.checkerboard.download(np.float32, (H // 2, W // 2))
        idx = rec.index.download(np.float32, (H // 4, W // 4, 2))
        render = rec.render_texture.download(np.uint16, (H // 4, W // 4, 4))
        dist = rec.distance.download(np.float32, (H // 4, W // 4))
        got = rec.reconstruct[1].download(np.uint16, (H // 2, W // 2, 4))  # swapped after the frame
        _same_bits(idx, oracle.cloud_shading_index(c, cb, "f32"), f"frame {frame} shading index")
        _same_bits(got, oracle.cloud_reconstruct(c, cb, idx, render, dist, history, W, H, "f32"),
                   f"frame {frame} reconstruction")
        _same_bits(d_hdr.download(np.uint16, (H, W, 4)), oracle.cloud_upscale(c, cb, depth, got, hdr0, "f32"),
                   f"frame {frame} composite")
        rgba = render.view(np.float16).astype(np.float32)
        print(f"frame {frame}: render T range {rgba[..., 3].min():.3f}-{rgba[..., 3].max():.3f}, "
              f"luminance max {rgba[..., :3].max():.3f}")
        assert np.isfinite(rgba).all() and rgba[..., 3].min() < 0.99
