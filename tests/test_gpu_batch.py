"""Batched LUT builds (skyhip_lut_batch_*) against the single-set path and the oracle.

Every batch item runs the single-set per-texel code, so each set's transmittance and
multiple-scattering LUTs and each view's sky-view LUT must equal, bit for bit, what
skyhip_transmittance_lut / skyhip_multiscattering_lut / skyhip_sky_view_lut produce for the
same blocks; a sample of items is also checked against the FP64 oracle (1e-3 relative).
Launch-shape changes (tuning, autotuning) must not change a single bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from skyrendering_amd import _abi, presets
from skyrendering_amd.batch import LutBatch

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _small_workload():
    sets, views = presets.batched_lut_workload(n_sets=6, suns_per_set=2)
    # Cover every size class explicitly.
    sizes_t = [(128, 32), (256, 64), (128, 32), (256, 64), (128, 32), (256, 64)]
    sizes_sv = [(96, 54), (192, 108), (384, 216), (96, 54), (192, 108), (384, 216)]
    sets = [(a, sizes_t[i], ms) for i, (a, _, ms) in enumerate(sets)]
    views = [(s, r, sizes_sv[s]) for s, r, _ in views]
    return sets, views


def _single(ctx, atm, t_size, ms_size):
    L = _abi.lib()
    tw, th = t_size
    mw, mh = ms_size
    T = ctx.alloc(tw * th * 16)
    M = ctx.alloc(mw * mh * 16)
    _abi.check(L.skyhip_transmittance_lut(ctx.handle, C.byref(atm), T.ptr, tw, th))
    _abi.check(L.skyhip_multiscattering_lut(ctx.handle, C.byref(atm), T.ptr, tw, th, M.ptr, mw, mh))
    return T, M


def _single_sv(ctx, atm, render, T, t_size, M, ms_size, sv_size):
    sw, sh = sv_size
    Lb, Tb = ctx.alloc(sw * sh * 16), ctx.alloc(sw * sh * 16)
    _abi.check(_abi.lib().skyhip_sky_view_lut(ctx.handle, C.byref(atm), C.byref(render), T.ptr, t_size[0],
                                              t_size[1], M.ptr, ms_size[0], ms_size[1], None, 1, Lb.ptr, Tb.ptr,
                                              sw, sh))
    return Lb.download(np.float32, (sh, sw, 4)), Tb.download(np.float32, (sh, sw, 4))


def test_batch_items_equal_single_set_path(ctx):
    sets, views = _small_workload()
    b = LutBatch(ctx, sets, views)
    b.launch()
    ctx.synchronize()
    singles = []
    for s, (atm, t_size, ms_size) in enumerate(sets):
        T, M = _single(ctx, atm, t_size, ms_size)
        bt, bm = b.set_outputs(s)
        assert np.array_equal(bt, T.download(np.float32, (t_size[1], t_size[0], 4))), f"T of set {s}"
        assert np.array_equal(bm, M.download(np.float32, (ms_size[1], ms_size[0], 4))), f"MS of set {s}"
        singles.append((T, M))
    for v, (s, render, sv_size) in enumerate(views):
        atm, t_size, ms_size = sets[s]
        T, M = singles[s]
        L, Tr = _single_sv(ctx, atm, render, T, t_size, M, ms_size, sv_size)
        bl, btr = b.view_outputs(v)
        assert np.array_equal(bl, L) and np.array_equal(btr, Tr), f"sky view of view {v}"
    steps = sum(t[0] * t[1] * 40 + 32 * 32 * 64 * 30 for _, t, _ in sets) + sum(
        sv[0] * sv[1] * 30 for _, _, sv in views)
    assert b.integrator_steps == steps


def test_batch_sample_against_oracle(ctx):
    sets, views = _small_workload()
    b = LutBatch(ctx, sets, views)
    b.launch()
    ctx.synchronize()
    for v in (1, 4, 11):
        s = views[v][0]
        atm, (tw, th), (mw, mh) = sets[s]
        bt, bm = b.set_outputs(s)
        eT = rel_err(bt[..., :3], oracle.transmittance_lut(atm, tw, th)[..., :3])[0]
        eM = rel_err(bm[..., :3], oracle.multiscattering_lut(atm, bt, mw, mh)[..., :3])[0]
        sw, sh = views[v][2]
        refL, refT = oracle.sky_view_lut(atm, views[v][1], bt, bm, sw, sh)
        bl, btr = b.view_outputs(v)
        eL = rel_err(bl[..., :3], refL[..., :3])[0]
        eTr = rel_err(btr[..., :3], refT[..., :3])[0]
        print(f"view {v} (set {s}, R_b {atm.bottom_radius:.0f} km, SV {sw}x{sh}): T {eT:.2e} MS {eM:.2e} "
              f"SV L {eL:.2e} SV T {eTr:.2e}")
        assert max(eT, eM, eL, eTr) < 1e-3


def test_batch_tuning_does_not_change_results(ctx):
    sets, views = _small_workload()
    b = LutBatch(ctx, sets, views)
    b.launch()
    ctx.synchronize()
    ref = [b.view_outputs(v)[0] for v in range(len(views))] + [b.set_outputs(s)[1] for s in range(len(sets))]
    b.set_tuning(transmittance_block=64, transmittance_tile=256, multiscattering_waves=1, sky_view_block=128,
                 sky_view_tile=512, sky_view_blocks_per_cu=2, sky_view_largest_first=0)
    b.launch()
    b.launch()  # relaunching the graph recomputes the same values
    ctx.synchronize()
    got = [b.view_outputs(v)[0] for v in range(len(views))] + [b.set_outputs(s)[1] for s in range(len(sets))]
    assert all(np.array_equal(x, y) for x, y in zip(ref, got))
    best, ms = b.autotune(repetitions=2)
    print("autotuned:", best, f"{ms:.3f} ms")
    assert ms > 0
    b.launch()
    ctx.synchronize()
    got = [b.view_outputs(v)[0] for v in range(len(views))] + [b.set_outputs(s)[1] for s in range(len(sets))]
    assert all(np.array_equal(x, y) for x, y in zip(ref, got))


def test_batch_rejects_bad_input(ctx):
    sets, views = _small_workload()
    bad_views = [(99, views[0][1], views[0][2])]
    with pytest.raises(_abi.SkyhipError):
        LutBatch(ctx, sets, bad_views)
    bad_sets = [(sets[0][0], (1, 32), (32, 32))]
    with pytest.raises(_abi.SkyhipError):
        LutBatch(ctx, bad_sets, [])
    b = LutBatch(ctx, sets, views)
    b.launch()
    ctx.synchronize()
    ref = [b.view_outputs(v)[0] for v in range(len(views))]
    with pytest.raises(_abi.SkyhipError):
        b.set_tuning(sky_view_block=96)
    # A rejected tuning leaves the batch launchable with its previous tuning and results.
    b.launch()
    ctx.synchronize()
    assert all(np.array_equal(x, b.view_outputs(v)[0]) for v, x in enumerate(ref))
