"""One frame's atmosphere LUT set (transmittance, multiple scattering, sky view, aerial
perspective) built by a single C call: skyhip_atmosphere_lut_set."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .atmosphere import AtmosphereParameters, AtmosphereRenderParameters
from .device import Context

_abi.register("skyhip_atmosphere_lut_set", C.c_int, [
    C.c_void_p, C.POINTER(_abi.AtmosphereBufferData), C.POINTER(_abi.AtmosphereRenderBufferData), C.c_void_p,
    C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
    C.c_void_p, C.c_int32, C.c_int32, C.c_int32])


class LutSet:
    def __init__(self, ctx: Context, sizes: dict):
        self.ctx = ctx
        self.sizes = sizes
        tw, th = sizes["transmittance"]
        mw, mh = sizes["multiscattering"]
        sw, sh = sizes["sky_view"]
        aw, ah, ad = sizes["aerial_perspective"]
        self.T = ctx.alloc(tw * th * 16)
        self.MS = ctx.alloc(mw * mh * 16)
        self.SV_L = ctx.alloc(sw * sh * 16)
        self.SV_T = ctx.alloc(sw * sh * 16)
        self.AP_L = ctx.alloc(aw * ah * ad * 16)
        self.AP_T = ctx.alloc(aw * ah * ad * 16)
        self._atm_key = None
        self.atm_block = None
        self.render_block = None

    def build(self, atmosphere: AtmosphereParameters, render: AtmosphereRenderParameters):
        key = id(atmosphere)
        if key != self._atm_key:
            self.atm_block = atmosphere.buffer_data()
            self._atm_key = key
        self.render_block = render.buffer_data(atmosphere.bottom_radius)
        s = self.sizes
        _abi.check(_abi.lib().skyhip_atmosphere_lut_set(
            self.ctx.handle, C.byref(self.atm_block), C.byref(self.render_block), self.T.ptr,
            *s["transmittance"], self.MS.ptr, *s["multiscattering"], self.SV_L.ptr, self.SV_T.ptr, *s["sky_view"],
            self.AP_L.ptr, self.AP_T.ptr, *s["aerial_perspective"]))

    def download(self):
        s = self.sizes
        (tw, th), (mw, mh), (sw, sh), (aw, ah, ad) = (s["transmittance"], s["multiscattering"], s["sky_view"],
                                                      s["aerial_perspective"])
        return {"T": self.T.download(np.float32, (th, tw, 4)), "MS": self.MS.download(np.float32, (mh, mw, 4)),
                "SV_L": self.SV_L.download(np.float32, (sh, sw, 4)), "SV_T": self.SV_T.download(np.float32, (sh, sw, 4)),
                "AP_L": self.AP_L.download(np.float32, (ad, ah, aw, 4)),
                "AP_T": self.AP_T.download(np.float32, (ad, ah, aw, 4))}
