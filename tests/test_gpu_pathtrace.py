"""GPU parity of the voxel-cloud path tracer (skyhip_cloud_path_trace) against the oracle.

- Grid generator: FP16 values within one unit in the last place of the oracle's, or 1e-6
  absolute where the threshold subtraction cancels (FP32 transcendental differences), in
  at most 1e-3 of the texels; majorants bit-exact.
- Single-majorant mode draws the reference's random sequence, so samples are compared
  one by one with the FP64 oracle: a sample matches when every channel agrees within
  1e-3 relative (floor 1e-6 of the image maximum).  A path whose FP32 and FP64
  arithmetic take a different discrete branch (a tentative collision accepted on one
  side only) diverges from there on; at least 90 % of the samples must match, and the
  mismatches must not move the mean (3 sigma).
- Brick-majorant mode (the workload's tracker) is compared sample by sample with the
  oracle's brick-majorant tracking (same random sequence, same bar as above), and its
  converged mean with the reference's single-majorant tracking, which draws different
  random numbers: with the FP64 oracle's (1024 vs 256 spp) through the image-sum z-score
  (within 3) and the coverage; per pixel, at 3072 spp each, with the single-majorant mode
  (batch means of 64 samples): image z within 3, at most 1 % of the pixels with |z| > 4
  (per-pixel luminance is heavy-tailed).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from skyrendering_amd import _abi
from skyrendering_amd.pathtrace import PathTraceScene, PathTracingInitParam, VoxelGrid, VoxelGridCreateInfo, path_trace

import _pt_cases as cases

pytestmark = pytest.mark.gpu


def _oracle_inputs(scene, bricks=False):
    T, sv_l, sv_t, ap_l, ap_t, grid = scene.host_inputs()
    majorants = scene.grid.download_majorants() if bricks else None
    return oracle.HostPathTraceInputs(scene.material, T, sv_l, sv_t, ap_l, ap_t, grid=grid, majorants=majorants,
                                      brick=scene.grid.brick if bricks else None)


def _gpu_samples(scene, frames, first=1):
    """Per-sample outputs [pixels, frames, 4] (accumulator cleared before every frame)."""
    n = scene.width * scene.height
    acc = scene.ctx.alloc(n * 16)
    out = np.empty((n, frames, 4), np.float32)
    for f in range(frames):
        acc.zero()
        path_trace(scene.ctx, scene.atmosphere.buffer, scene.renderer.buffer, scene.common, scene.material,
                   scene.tracer.params, scene.inputs, (0, 0, scene.width, scene.height), first + f, 1, acc, None,
                   scene.width, scene.height)
        out[:, f] = acc.download(np.float32, (n, 4))
    acc.free()
    return out


def _gpu_batch_means(scene, batches, per_batch, first=1):
    """Batch means [pixels, batches, 4] of `per_batch` frames each (one launch per batch)."""
    n = scene.width * scene.height
    acc = scene.ctx.alloc(n * 16)
    out = np.empty((n, batches, 4), np.float64)
    for b in range(batches):
        acc.zero()
        path_trace(scene.ctx, scene.atmosphere.buffer, scene.renderer.buffer, scene.common, scene.material,
                   scene.tracer.params, scene.inputs, (0, 0, scene.width, scene.height), first + b * per_batch,
                   per_batch, acc, None, scene.width, scene.height)
        out[:, b] = acc.download(np.float32, (n, 4)) / per_batch
    acc.free()
    return out


def test_voxel_grid_generator_matches_oracle(ctx):
    ci = VoxelGridCreateInfo()
    dims = (64, 48, 40)
    g = VoxelGrid(ctx, dims).generate(ci)
    got = g.download()
    ref = oracle.voxel_grid_f16(ci.to_c(), dims)
    ulps = np.abs(got.astype(np.int32) - ref.astype(np.int32))
    gv, rv = got.view(np.float16).astype(np.float64), ref.view(np.float16).astype(np.float64)
    # exp(.) * falloff - threshold cancels near the threshold, where FP32 rounding (~1e-7
    # absolute) spans several FP16 ulps of the small result: allow 1 ulp or 1e-6 absolute.
    ok = (ulps <= 1) | (np.abs(gv - rv) <= 1e-6)
    frac = float(np.mean(ulps > 0))
    print(f"grid {dims}: max ulp {ulps.max()}, max abs {np.abs(gv - rv).max():.2e}, differing texels {frac:.2e}, "
          f"nonzero {np.mean(ref > 0):.3f}")
    assert np.all(ok) and frac <= 1e-3
    assert np.array_equal(g.download_majorants(), oracle.brick_majorants(got))
    for brick in (8, 32):
        gb = VoxelGrid(ctx, dims, brick=brick).upload(got)
        assert np.array_equal(gb.download_majorants(), oracle.brick_majorants(got, brick)), brick


def _cells_reference(grid):
    """Cell-corner copy built in numpy: cell (i, j, k) from (-1, -1, -1), corners a + 2b + 4c."""
    dz, dy, dx = grid.shape
    pad = np.zeros((dz + 2, dy + 2, dx + 2), np.uint16)
    pad[1:-1, 1:-1, 1:-1] = grid
    out = np.empty((dz + 1, dy + 1, dx + 1, 8), np.uint16)
    for k in range(8):
        a, b, c = k & 1, (k >> 1) & 1, k >> 2
        out[..., k] = pad[c:c + dz + 1, b:b + dy + 1, a:a + dx + 1]
    return out


def test_cell_corner_copy_and_identical_samples(ctx):
    """skyhip_voxel_grid_cells equals the numpy construction (including the zero border) on a
    ragged grid, and brick-majorant tracking gives bit-identical samples with and without it."""
    dims = (37, 29, 23)
    g = VoxelGrid(ctx, dims, brick=8).generate(VoxelGridCreateInfo())
    assert np.array_equal(g.download_cells(), _cells_reference(g.download()))
    w = cases.small_grid_workload()
    with_cells = PathTraceScene(ctx, 32, 32, workload=w, bricks=True)
    assert with_cells.inputs.cell_corners
    plain_grid = VoxelGrid(ctx, with_cells.grid.dims, brick=with_cells.grid.brick, cells=False)
    plain_grid.upload(with_cells.grid.download())
    without = PathTraceScene(ctx, 32, 32, workload=cases.small_grid_workload(), grid=plain_grid, bricks=True)
    assert not without.inputs.cell_corners
    a, b = _gpu_samples(with_cells, 3), _gpu_samples(without, 3)
    print(f"cell-corner lookups: {np.mean(a[..., 3] == 0):.3f} of samples scattered; identical: {np.array_equal(a, b)}")
    assert np.array_equal(a, b)


def test_refill_kernel_matches_per_pixel_kernel(ctx):
    """With 4^3 bricks and the cell-corner copy bound, brick-majorant frames run on the persistent
    kernel that refills finished lanes after every bounce, with nested tracking loops; without
    the copy, on the one-lane-per-pixel kernel with flat tracking loops.  A three-frame call over
    a region off the origin whose sides are not multiples of the 8x8 tiles gives bit-identical
    accumulators and the same rendered mask, and pixels outside the region stay untouched."""
    w = cases.small_grid_workload()
    w["brick"] = 4
    refill = PathTraceScene(ctx, 30, 22, workload=w, bricks=True)
    assert refill.inputs.cell_corners and refill.grid.brick == 4
    plain_grid = VoxelGrid(ctx, refill.grid.dims, brick=4, cells=False)
    plain_grid.upload(refill.grid.download())
    w2 = cases.small_grid_workload()
    w2["brick"] = 4
    per_pixel = PathTraceScene(ctx, 30, 22, workload=w2, grid=plain_grid, bricks=True)
    assert not per_pixel.inputs.cell_corners
    region = (3, 5, 27, 19)
    out = []
    for scene in (refill, per_pixel):
        n = scene.width * scene.height
        seed = np.full((n, 4), 0.25, np.float32)
        acc = ctx.upload(seed)
        mask = ctx.upload(np.zeros(n, np.uint8))
        path_trace(ctx, scene.atmosphere.buffer, scene.renderer.buffer, scene.common, scene.material,
                   scene.tracer.params, scene.inputs, region, 7, 3, acc, mask, scene.width, scene.height)
        out.append((acc.download(np.float32, (scene.height, scene.width, 4)),
                    mask.download(np.uint8, (scene.height, scene.width))))
        acc.free()
        mask.free()
    (a, ma), (b, mb) = out
    inside = np.zeros((22, 30), bool)
    inside[5:19, 3:27] = True
    print(f"refill vs per-pixel: identical {np.array_equal(a, b)}, "
          f"{np.mean(a[inside][:, 3] < 3.25):.3f} of region pixels scattered in some frame")
    assert np.array_equal(a, b) and np.array_equal(ma, mb)
    assert np.all(ma[inside] == 1) and np.all(ma[~inside] == 0)
    assert np.all(a[~inside] == 0.25)


def test_single_majorant_samples_match_oracle(ctx):
    w = cases.small_grid_workload()
    scene = PathTraceScene(ctx, 32, 32, workload=w, bricks=False)
    frames = 4
    gpu = _gpu_samples(scene, frames)
    ref = oracle.path_trace(scene.atmosphere.buffer, scene.renderer.buffer, scene.common, scene.tracer.params,
                            _oracle_inputs(scene), 32, 32, np.arange(32 * 32), 1, frames)
    floor = 1e-6 * np.abs(ref[..., :3]).max()
    err = np.abs(gpu - ref) / np.maximum(np.abs(ref), floor)
    match = np.all(err <= 1e-3, axis=-1)
    scattered = ref[..., 3] == 0
    frac = float(match.mean())
    lum_g, lum_o = gpu[..., :3].sum(-1).ravel(), ref[..., :3].sum(-1).ravel()
    se = math.sqrt(lum_g.var() / lum_g.size + lum_o.var() / lum_o.size)
    z = (lum_g.mean() - lum_o.mean()) / se
    print(f"single majorant: {frac:.4f} of samples match ({scattered.mean():.3f} scattered, "
          f"{match[scattered].mean():.4f} of those); mean z {z:.2f}")
    assert scattered.mean() > 0.1
    assert frac >= 0.9 and abs(z) < 3.0


@pytest.mark.parametrize("brick", [16, 8])
def test_brick_majorant_samples_match_oracle(ctx, brick):
    w = cases.small_grid_workload()
    w["brick"] = brick
    scene = PathTraceScene(ctx, 32, 32, workload=w, bricks=True)
    assert scene.grid.brick == brick
    frames = 4
    gpu = _gpu_samples(scene, frames)
    ref = oracle.path_trace(scene.atmosphere.buffer, scene.renderer.buffer, scene.common, scene.tracer.params,
                            _oracle_inputs(scene, bricks=True), 32, 32, np.arange(32 * 32), 1, frames)
    floor = 1e-6 * np.abs(ref[..., :3]).max()
    err = np.abs(gpu - ref) / np.maximum(np.abs(ref), floor)
    match = np.all(err <= 1e-3, axis=-1)
    scattered = ref[..., 3] == 0
    frac = float(match.mean())
    lum_g, lum_o = gpu[..., :3].sum(-1).ravel(), ref[..., :3].sum(-1).ravel()
    se = math.sqrt(lum_g.var() / lum_g.size + lum_o.var() / lum_o.size)
    z = (lum_g.mean() - lum_o.mean()) / se
    print(f"{brick}^3 brick majorants: {frac:.4f} of samples match ({scattered.mean():.3f} scattered, "
          f"{match[scattered].mean():.4f} of those); mean z {z:.2f}")
    assert scattered.mean() > 0.1
    assert frac >= 0.9 and abs(z) < 3.0


def _z_stats(bm_a, bm_b):
    """Per-pixel and image-sum z-scores of two sets of batch means [pix, batches] (luminance)."""
    ma, va = bm_a.mean(1), bm_a.var(1, ddof=1) / bm_a.shape[1]
    mb, vb = bm_b.mean(1), bm_b.var(1, ddof=1) / bm_b.shape[1]
    z_img = (ma.sum() - mb.sum()) / math.sqrt(va.sum() + vb.sum())
    live = (va + vb) > 0
    z_pix = (ma[live] - mb[live]) / np.sqrt(va[live] + vb[live])
    return z_img, z_pix, ma, mb


def test_brick_majorants_converge_to_reference_mean(ctx):
    w = cases.small_grid_workload()
    scene = PathTraceScene(ctx, 32, 32, workload=w, bricks=True)
    # (1) Against the FP64 oracle directly: image-sum mean and coverage.
    bm = _gpu_batch_means(scene, batches=32, per_batch=32)  # 1024 spp
    ref = oracle.path_trace(scene.atmosphere.buffer, scene.renderer.buffer, scene.common, scene.tracer.params,
                            _oracle_inputs(scene), 32, 32, np.arange(32 * 32), 100001, 256)
    o = ref[..., :3].sum(-1)
    ob = o.reshape(o.shape[0], 16, 16).mean(-1)  # 16 batches of 16 samples
    z_img, _, mg, mo = _z_stats(bm[..., :3].sum(-1), ob)
    a_g, a_o = bm[..., 3].mean(), ref[..., 3].mean()
    print(f"bricks vs oracle: image z {z_img:.2f}, mean {mg.mean():.4e} vs {mo.mean():.4e}, "
          f"alpha {a_g:.4f} vs {a_o:.4f}")
    assert abs(z_img) < 3.0
    assert abs(a_g - a_o) < 4 * math.sqrt(a_o * (1 - a_o) / ref[..., 3].size) + 1e-3
    # (2) Per pixel, at 3072 spp, against the single-majorant tracker (whose samples are the
    # oracle's one for one, test above): batch means of 64 samples.
    ref_scene = PathTraceScene(ctx, 32, 32, workload=cases.small_grid_workload(), grid=scene.grid, bricks=False)
    bb = _gpu_batch_means(scene, batches=48, per_batch=64, first=1)[..., :3].sum(-1)
    rb = _gpu_batch_means(ref_scene, batches=48, per_batch=64, first=500001)[..., :3].sum(-1)
    z_img, z_pix, mb_, mr_ = _z_stats(bb, rb)
    frac = float(np.mean(np.abs(z_pix) > 4.0))
    print(f"bricks vs single majorant (3072 spp): image z {z_img:.2f}, pixels |z| > 4: {frac:.4f} of {z_pix.size}, "
          f"max |z| {np.abs(z_pix).max():.2f}")
    assert abs(z_img) < 3.0 and frac <= 0.01


def test_tiles_frames_and_display(ctx):
    w = cases.small_grid_workload()
    w["init"] = PathTracingInitParam(**{**w["init"].__dict__, "sqrt_tile_count": 2})
    scene = PathTraceScene(ctx, 30, 22, workload=w)
    pt = scene.tracer
    regions = []
    for _ in range(4):  # one frame in four tiles (uneven split: 30 = 15 + 15, 22 = 11 + 11)
        regions.append(pt.get_render_region())
        pt.render()
    assert regions == [(0, 0, 15, 11), (15, 0, 30, 11), (0, 11, 15, 22), (15, 11, 30, 22)]
    tiled = pt.accumulated()
    one = PathTraceScene(ctx, 30, 22, workload=cases.small_grid_workload(), grid=scene.grid)
    one.tracer.render_frames(1)
    assert np.array_equal(tiled, one.tracer.accumulated())
    # Display pass: color.rgb = color.rgb * avg.a + avg.rgb, avg = accum / frame.
    one.tracer.render_frames(2)
    rng = np.random.default_rng(3)
    disp_in = rng.random((22, 30, 4)).astype(np.float16)
    disp = ctx.upload(disp_in)
    one.tracer.display(disp)
    got = disp.download(np.float16, (22, 30, 4)).astype(np.float32)
    avg = one.tracer.accumulated() / 3.0
    want = disp_in.astype(np.float32)
    want[..., :3] = want[..., :3] * avg[..., 3:4] + avg[..., :3]
    want16 = want.astype(np.float16).astype(np.float32)
    ulp = np.abs(got - want16) / np.maximum(np.spacing(np.abs(want16).astype(np.float16)).astype(np.float32), 1e-8)
    assert np.all(ulp <= 1.0) and np.array_equal(got[..., 3], disp_in[..., 3].astype(np.float32))


def test_counting_variant_and_rejections(ctx):
    w = cases.small_grid_workload()
    scene = PathTraceScene(ctx, 24, 24, workload=w)
    c = scene.count_work(frames=2)
    print("work:", c)
    assert c["samples"] == 24 * 24 * 2 and c["lookups"] > 0 and c["segments"] > 0 and c["vertices"] > 0
    assert c["shadow_rays"] >= c["vertices"] * 0.5
    # Counting and timing variants produce the same accumulation.
    acc = ctx.alloc(24 * 24 * 16)
    acc.zero()
    path_trace(ctx, scene.atmosphere.buffer, scene.renderer.buffer, scene.common, scene.material, scene.tracer.params,
               scene.inputs, (0, 0, 24, 24), 1, 2, acc, None, 24, 24, scene.counter)
    a1 = acc.download(np.float32, (24, 24, 4))
    acc.zero()
    path_trace(ctx, scene.atmosphere.buffer, scene.renderer.buffer, scene.common, scene.material, scene.tracer.params,
               scene.inputs, (0, 0, 24, 24), 1, 2, acc, None, 24, 24)
    assert np.array_equal(a1, acc.download(np.float32, (24, 24, 4)))
    L = _abi.lib()
    reg = (C.c_int32 * 4)(0, 0, 25, 24)
    args = lambda r, fr, prm, mat, inp: (ctx.handle, C.byref(scene.atmosphere.buffer), C.byref(scene.renderer.buffer),
                                         C.byref(scene.common), C.byref(mat), C.byref(prm), C.byref(inp), r, 1, fr,
                                         acc.ptr, None, 24, 24, None)
    assert L.skyhip_cloud_path_trace(*args(reg, 1, scene.tracer.params, scene.material, scene.inputs)) == \
        _abi.ERR_INVALID_SIZE
    ok = (C.c_int32 * 4)(0, 0, 24, 24)
    assert L.skyhip_cloud_path_trace(*args(ok, -1, scene.tracer.params, scene.material, scene.inputs)) == \
        _abi.ERR_INVALID_SIZE
    bad = _abi.PathTracingParametersC()
    C.memmove(C.byref(bad), C.byref(scene.tracer.params), C.sizeof(bad))
    bad.prng = 7
    assert L.skyhip_cloud_path_trace(*args(ok, 1, bad, scene.material, scene.inputs)) == _abi.ERR_INVALID_ARGUMENT
    minimal = cases.minimal_material(1.0)
    assert L.skyhip_cloud_path_trace(*args(ok, 1, scene.tracer.params, minimal, scene.inputs)) == \
        _abi.ERR_INVALID_ARGUMENT  # brick majorants need a VOXEL_F16 material
    # A grid of more than 2^31 cells has no cell-corner copy (skyhip_voxel_grid_cells refuses it):
    # bound with one, the call is refused.  Zero frames, so that nothing would be launched even if
    # the check let it through.
    tall = _abi.CloudMaterial()
    C.memmove(C.byref(tall), C.byref(scene.material), C.sizeof(tall))
    tall.voxel_grid_dim[0], tall.voxel_grid_dim[1], tall.voxel_grid_dim[2] = 2048, 2048, 2048
    assert scene.inputs.cell_corners
    assert L.skyhip_cloud_path_trace(*args(ok, 0, scene.tracer.params, tall, scene.inputs)) == _abi.ERR_INVALID_SIZE
    assert b"cell_corners" in L.skyhip_last_error()
    acc.free()


def test_bound_froxel_samples_match_oracle(ctx):
    """With a ray-scatter-visibility froxel bound (skyhip_bind_shadow_froxel), each scattered
    sample's aerial-perspective luminance is scaled by SampleRayScatterVisibility
    (VolumetricCloudPathTracing.comp:277): per-sample parity against the oracle with the same
    froxel bound, and the samples differ from the unbound ones."""
    from skyrendering_amd.cloud import bind_shadow_froxel
    w = cases.small_grid_workload()
    scene = PathTraceScene(ctx, 32, 32, workload=w, bricks=False)
    frames = 2
    unbound = _gpu_samples(scene, frames)
    fw, fh, fd = 8, 6, 16
    rng = np.random.default_rng(12)
    froxel = (65535 * (0.2 + 0.8 * rng.random((fd, fh, fw)))).astype(np.uint16)
    d_froxel = ctx.upload(froxel)
    bind_shadow_froxel(ctx, d_froxel, (fw, fh, fd))
    oracle.bind_shadow_froxel(froxel)
    try:
        gpu = _gpu_samples(scene, frames)
        ref = oracle.path_trace(scene.atmosphere.buffer, scene.renderer.buffer, scene.common, scene.tracer.params,
                                _oracle_inputs(scene), 32, 32, np.arange(32 * 32), 1, frames)
    finally:
        bind_shadow_froxel(ctx, None)
        oracle.bind_shadow_froxel(None)
    floor = 1e-6 * np.abs(ref[..., :3]).max()
    err = np.abs(gpu - ref) / np.maximum(np.abs(ref), floor)
    match = np.all(err <= 1e-3, axis=-1)
    scattered = ref[..., 3] == 0
    changed = np.any(np.abs(gpu - unbound) > 1e-6 * np.abs(unbound).max(), axis=-1)
    print(f"froxel bound: {match.mean():.4f} of samples match; {changed[scattered].mean():.3f} of scattered samples "
          f"changed by the froxel")
    assert scattered.mean() > 0.1 and match.mean() >= 0.9
    assert changed[scattered].mean() > 0.5 and not changed[~scattered].any()
