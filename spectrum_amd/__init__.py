"""spectrum_amd: MI355X-native batched FFT -> |X| -> spectrum-maxima path.

Layers (bottom up):
  lib/libspectrum_mi355x.so  HIP kernels for gfx950 + the C ABI (include/spectrum_mi355x.h)
  capi                       ctypes binding of that ABI (Context, DeviceBuffer, siggen)
  lib/dsp_spectrum*.so       pybind11 module mirroring the reference's Python surface
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
