"""Batched atmosphere LUT builds (skyhip_lut_batch_*): many atmosphere sets and (set, sun)
views at mixed LUT sizes, built by one hipGraph launch whose launch shape an autotuner
picks.  Each set's transmittance / multiple-scattering LUTs and each view's sky-view LUT
equal what the single-set calls (Atmosphere.update_luts, AtmosphereRenderer.render)
produce for the same blocks.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .atmosphere import AtmosphereParameters, AtmosphereRenderParameters
from .device import Context

_P = C.POINTER
_abi.register("skyhip_lut_batch_create", C.c_int, [C.c_void_p, _P(_abi.LutBatchSetC), C.c_int32,
                                                   _P(_abi.LutBatchViewC), C.c_int32, _P(C.c_void_p)])
_abi.register("skyhip_lut_batch_destroy", C.c_int, [C.c_void_p])
_abi.register("skyhip_lut_batch_launch", C.c_int, [C.c_void_p])
_abi.register("skyhip_lut_batch_get_tuning", C.c_int, [C.c_void_p, _P(_abi.LutBatchTuningC)])
_abi.register("skyhip_lut_batch_set_tuning", C.c_int, [C.c_void_p, _P(_abi.LutBatchTuningC)])
_abi.register("skyhip_lut_batch_autotune", C.c_int, [C.c_void_p, C.c_int32, _P(_abi.LutBatchTuningC),
                                                     _P(C.c_float)])
_abi.register("skyhip_lut_batch_set_outputs", C.c_int, [C.c_void_p, C.c_int32, _P(C.c_void_p), _P(C.c_void_p)])
_abi.register("skyhip_lut_batch_view_outputs", C.c_int, [C.c_void_p, C.c_int32, _P(C.c_void_p), _P(C.c_void_p)])
_abi.register("skyhip_lut_batch_integrator_steps", C.c_int, [C.c_void_p, _P(C.c_ulonglong)])


def tuning_dict(t: _abi.LutBatchTuningC) -> dict:
    return {name: int(getattr(t, name)) for name, _ in t._fields_}


class LutBatch:
    """sets: [(AtmosphereBufferData, (tw, th), (mw, mh))]; views: [(set index, render block, (sw, sh))]."""

    def __init__(self, ctx: Context, sets, views):
        self.handle = None
        self.ctx = ctx
        cs = (_abi.LutBatchSetC * len(sets))()
        self.set_sizes = []
        for i, (atm, t_size, ms_size) in enumerate(sets):
            cs[i].atmosphere = atm
            cs[i].transmittance_width, cs[i].transmittance_height = t_size
            cs[i].multiscattering_width, cs[i].multiscattering_height = ms_size
            self.set_sizes.append((tuple(t_size), tuple(ms_size)))
        cv = (_abi.LutBatchViewC * max(len(views), 1))()
        self.view_sizes, self.view_sets = [], []
        for i, (s, render, sv_size) in enumerate(views):
            cv[i].set = s
            cv[i].sky_view_width, cv[i].sky_view_height = sv_size
            cv[i].render = render
            self.view_sizes.append(tuple(sv_size))
            self.view_sets.append(s)
        self.sets_c, self.views_c = cs, cv
        h = C.c_void_p()
        _abi.check(_abi.lib().skyhip_lut_batch_create(ctx.handle, cs, len(sets), cv, len(views), C.byref(h)))
        self.handle = h

    def close(self):
        """Frees the batch; a batch whose context is already closed is left to process exit
        (its memory went with the context's device state)."""
        if self.handle and getattr(self.ctx, "handle", None) is not None:
            _abi.lib().skyhip_lut_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def launch(self):
        _abi.check(_abi.lib().skyhip_lut_batch_launch(self.handle))

    @property
    def tuning(self) -> dict:
        t = _abi.LutBatchTuningC()
        _abi.check(_abi.lib().skyhip_lut_batch_get_tuning(self.handle, C.byref(t)))
        return tuning_dict(t)

    def set_tuning(self, **kw):
        t = _abi.LutBatchTuningC()
        _abi.check(_abi.lib().skyhip_lut_batch_get_tuning(self.handle, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        _abi.check(_abi.lib().skyhip_lut_batch_set_tuning(self.handle, C.byref(t)))

    def autotune(self, repetitions=3):
        t = _abi.LutBatchTuningC()
        ms = C.c_float()
        _abi.check(_abi.lib().skyhip_lut_batch_autotune(self.handle, repetitions, C.byref(t), C.byref(ms)))
        return tuning_dict(t), ms.value

    @property
    def integrator_steps(self) -> int:
        n = C.c_ulonglong()
        _abi.check(_abi.lib().skyhip_lut_batch_integrator_steps(self.handle, C.byref(n)))
        return n.value

    def set_output_ptrs(self, s):
        """Device pointers (transmittance, multiscattering) of set s."""
        t, m = C.c_void_p(), C.c_void_p()
        _abi.check(_abi.lib().skyhip_lut_batch_set_outputs(self.handle, s, C.byref(t), C.byref(m)))
        return t.value, m.value

    def view_output_ptrs(self, v):
        """Device pointers (luminance, transmittance) of view v's sky-view LUTs."""
        l, t = C.c_void_p(), C.c_void_p()
        _abi.check(_abi.lib().skyhip_lut_batch_view_outputs(self.handle, v, C.byref(l), C.byref(t)))
        return l.value, t.value

    def set_outputs(self, s):
        """(transmittance, multiscattering) of set s as host arrays [h, w, 4]."""
        t, m = C.c_void_p(), C.c_void_p()
        _abi.check(_abi.lib().skyhip_lut_batch_set_outputs(self.handle, s, C.byref(t), C.byref(m)))
        (tw, th), (mw, mh) = self.set_sizes[s]
        return self.ctx.download_ptr(t.value, np.float32, (th, tw, 4)), self.ctx.download_ptr(m.value, np.float32,
                                                                                               (mh, mw, 4))

    def view_outputs(self, v):
        """(luminance, transmittance) sky-view LUTs of view v as host arrays [h, w, 4]."""
        l, t = C.c_void_p(), C.c_void_p()
        _abi.check(_abi.lib().skyhip_lut_batch_view_outputs(self.handle, v, C.byref(l), C.byref(t)))
        sw, sh = self.view_sizes[v]
        return self.ctx.download_ptr(l.value, np.float32, (sh, sw, 4)), self.ctx.download_ptr(t.value, np.float32,
                                                                                              (sh, sw, 4))
