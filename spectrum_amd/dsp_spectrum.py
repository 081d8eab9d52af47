"""Loads the in-tree pybind11 module ``dsp_spectrum`` (the reference's Python
surface, python/dsp_spectrum_module.cpp there) and registers it under its own
name, so ``from spectrum_amd import dsp_spectrum`` and, afterwards, a plain
``import dsp_spectrum`` both resolve to spectrum_amd/lib/dsp_spectrum*.so."""
import glob
import importlib.util
import os
import sys

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")


def _load():
    if "dsp_spectrum" in sys.modules and hasattr(sys.modules["dsp_spectrum"], "FFTProcessorROCm"):
        return sys.modules["dsp_spectrum"]
    found = sorted(glob.glob(os.path.join(_LIB, "dsp_spectrum*.so")))
    if not found:
        raise ImportError(f"dsp_spectrum extension not built under {_LIB}: run __graft_entry__.build() "
                          "or `make -C spectrum_amd/csrc`")
    spec = importlib.util.spec_from_file_location("dsp_spectrum", found[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    sys.modules["dsp_spectrum"] = mod
    return mod


_mod = _load()
globals().update({k: v for k, v in _mod.__dict__.items() if not k.startswith("__")})
__doc__ = _mod.__doc__
__file__ = _mod.__file__
