"""MI355X (gfx950) device path of the implicit 2D DG / condensed-HDG Euler solver.

The product is ``libdgh_device.so`` (HIP kernels behind the C ABI declared in
``include/dgh/dgh_device.h``) and the ``bin/dg_hdg_2d`` CLI. ``device`` is a
thin ctypes binding of that ABI for Python callers and the check harness.
"""
from .device import DeviceContext, DeviceContext1D, DeviceError, load_library  # noqa: F401

__all__ = ["DeviceContext", "DeviceContext1D", "DeviceError", "load_library"]
