"""skyrendering_amd: MI355X (gfx950) HIP implementation of SkyRendering's data-parallel passes.

The compute path is libskyhip.so (HIP kernels behind the C ABI in include/skyhip.h).  This
package is the Python view of that ABI, mirroring the reference's host classes
(Atmosphere, AtmosphereRenderer, VolumetricCloud, ...).
"""
from . import _abi
from .atmosphere import (Atmosphere, AtmosphereParameters, AtmosphereRenderer, AtmosphereRenderInitParameters,
                         AtmosphereRenderParameters, Camera)
from .device import Context, DeviceBuffer

__all__ = [
    "Atmosphere", "AtmosphereParameters", "AtmosphereRenderer", "AtmosphereRenderInitParameters",
    "AtmosphereRenderParameters", "Camera", "Context", "DeviceBuffer",
]
