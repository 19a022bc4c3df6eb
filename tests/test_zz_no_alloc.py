"""The allocation-free surrogate gradient (skyhip_surrogate_gradient_noalloc) against the
workspace entry point (skyhip_surrogate_gradient): both run the same training and reduction
kernels on the same inputs, so the gradients and losses must be bit-identical.  Two calls in a
row, the second with fewer samples (the shared device array keeps the first call's partials;
the second must not see them), for both networks, through the C ABI; plus the size limit."""
import ctypes as C

import numpy as np
import pytest

from skyrendering_amd import _abi
from skyrendering_amd import surrogate as sg

pytestmark = pytest.mark.gpu

_abi.register("skyhip_surrogate_gradient_noalloc", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                                             C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_void_p,
                                                             C.c_void_p])
MAX_COUNT = 8192  # SKYHIP_SURROGATE_NOALLOC_MAX_COUNT


def _inputs(ctx, n, seed):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 8), np.float32)
    f[:, :7] = rng.uniform(-1, 1, size=(n, 7))
    t = rng.normal(-2.0, 0.5, size=(n, 3)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
    return ctx.upload(f), ctx.upload(t), ctx.upload(w)


@pytest.mark.parametrize("network", ["f16_gate", "bf16_silu"])
def test_noalloc_gradient_equals_workspace_gradient(ctx, network):
    L = _abi.lib()
    params = ctx.upload((sg.init_params(11) * 2.0).astype(np.float32))
    nid = sg.NETWORKS[network]
    for n, seed in ((MAX_COUNT, 1), (1000, 2)):  # the second call is smaller than the first
        f, t, w = _inputs(ctx, n, seed)
        ws_bytes = L.skyhip_surrogate_workspace_bytes(n)
        ws = ctx.alloc(ws_bytes)
        g_ref, l_ref = ctx.alloc(13187 * 4), ctx.alloc(4)
        g_new, l_new = ctx.alloc(13187 * 4), ctx.alloc(4)
        _abi.check(L.skyhip_surrogate_gradient(ctx.handle, nid, params.ptr, f.ptr, t.ptr, w.ptr, n, 0.5, g_ref.ptr,
                                               l_ref.ptr, ws.ptr, ws_bytes))
        _abi.check(L.skyhip_surrogate_gradient_noalloc(ctx.handle, nid, params.ptr, f.ptr, t.ptr, w.ptr, n, 0.5,
                                                       g_new.ptr, l_new.ptr))
        ctx.synchronize()
        a, b = g_ref.download(np.float32, (13187,)), g_new.download(np.float32, (13187,))
        la, lb = l_ref.download(np.float32, (1,)), l_new.download(np.float32, (1,))
        print(f"{network} n={n}: |grad| {np.linalg.norm(a):.4e}, max |diff| {np.abs(a - b).max():.3e}, "
              f"loss {la[0]:.6e} / {lb[0]:.6e}")
        assert np.isfinite(a).all() and np.linalg.norm(a) > 0
        assert np.array_equal(a, b) and np.array_equal(la, lb)
        for buf in (ws, g_ref, l_ref, g_new, l_new, f, t, w):
            buf.free()


def test_noalloc_gradient_rejects_counts_above_the_maximum(ctx):
    L = _abi.lib()
    params = ctx.upload(sg.init_params(11).astype(np.float32))
    f, t, w = _inputs(ctx, 64, 3)
    g = ctx.alloc(13187 * 4)
    rc = L.skyhip_surrogate_gradient_noalloc(ctx.handle, 0, params.ptr, f.ptr, t.ptr, None, MAX_COUNT + 1, 1.0, g.ptr,
                                             None)
    assert rc == _abi.ERR_INVALID_SIZE
    assert b"allocation-free maximum" in L.skyhip_last_error()
