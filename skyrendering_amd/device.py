"""Device context and buffers over the skyhip C ABI (no torch dependency)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


class DeviceBuffer:
    """A device allocation owned by a Context (freed explicitly or on GC)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _abi.check(_abi.lib().skyhip_malloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def free(self):
        if self.ptr is not None and self.ctx.handle is not None:
            _abi.check(_abi.lib().skyhip_free(self.ctx.handle, self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, array: np.ndarray):
        a = np.ascontiguousarray(array)
        if a.nbytes > self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into a {self.nbytes}-byte buffer")
        _abi.check(_abi.lib().skyhip_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes))
        return self

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError(f"download of {out.nbytes} bytes from a {self.nbytes}-byte buffer")
        _abi.check(_abi.lib().skyhip_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self):
        _abi.check(_abi.lib().skyhip_memset(self.ctx.handle, self.ptr, 0, self.nbytes))
        return self


class Context:
    """One HIP device and one stream (SkyhipContext)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _abi.check(_abi.lib().skyhip_create(device, C.byref(h)))
        self.handle = h.value

    def close(self):
        if self.handle is not None:
            _abi.lib().skyhip_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def upload(self, array: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(array)
        return DeviceBuffer(self, max(a.nbytes, 1)).upload(a)

    def download_ptr(self, ptr: int, dtype, shape) -> np.ndarray:
        """Copies `shape` elements from a device pointer the library handed out."""
        out = np.empty(shape, dtype=dtype)
        _abi.check(_abi.lib().skyhip_memcpy_d2h(self.handle, out.ctypes.data, ptr, out.nbytes))
        return out

    def synchronize(self):
        _abi.check(_abi.lib().skyhip_synchronize(self.handle))

    def record(self, slot: int):
        _abi.check(_abi.lib().skyhip_event_record(self.handle, slot))

    def elapsed_ms(self, start_slot: int, end_slot: int) -> float:
        ms = C.c_float()
        _abi.check(_abi.lib().skyhip_event_elapsed_ms(self.handle, start_slot, end_slot, C.byref(ms)))
        return float(ms.value)

    def device_info(self) -> dict:
        cu, wave, clk = C.c_int32(), C.c_int32(), C.c_int32()
        name = C.create_string_buffer(256)
        _abi.check(_abi.lib().skyhip_device_info(self.handle, C.byref(cu), C.byref(wave), C.byref(clk), name, 256))
        return {"cu_count": cu.value, "wave_size": wave.value, "clock_khz": clk.value, "arch": name.value.decode()}
