"""CPU-side checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/spectrum_mi355x.h declares, the ctypes table matches
the header, records keep the reference's 32-byte MaxValue layout, and the host
generator is deterministic.  No compute calls that need a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from spectrum_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spectrum_mi355x.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(spx_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_what_ctypes_binds():
    assert header_functions() == sorted(capi.exported_symbols())


def test_library_exports_every_symbol():
    lib = capi.load()
    for name in header_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(spx_\w+)", out))
    assert set(header_functions()) <= exported


def test_version_string():
    assert b"gfx950" in capi.load().spx_version()


def test_maxvalue_layout_matches_reference():
    # include/spectrum/types/spectrum_result_types.hpp:34-43: u32, 6 x f32, u32 = 32 bytes
    assert capi.MAXVALUE_DTYPE.itemsize == 32
    assert [capi.MAXVALUE_DTYPE.fields[f][1] for f in capi.MAXVALUE_DTYPE.names] == [0, 4, 8, 12, 16, 20, 24, 28]


def test_fused_desc_layout():
    assert ctypes.sizeof(capi.SpxFusedDesc) == 8 + 8 + 4 * 12 + 8 * 4


def test_create_without_gpu_reports_error():
    lib = capi.load()
    h = ctypes.c_void_p()
    n = ctypes.c_int()
    # in a GPU-less container spx_create must fail cleanly (no fallback)
    import torch  # noqa: F401 (torch only tells us whether a device exists)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    assert lib.spx_create(0, ctypes.byref(h)) != capi.SPX_OK


def test_siggen_deterministic_and_subsettable():
    a = capi.siggen(capi.GEN_MULTI_TONE, 7, 6, 1024)
    b = capi.siggen(capi.GEN_MULTI_TONE, 7, 6, 1024, threads=1)
    assert np.array_equal(a, b)
    c = capi.siggen(capi.GEN_MULTI_TONE, 7, 2, 1024, chan_offset=3)
    assert np.array_equal(a[3:5], c)
    d = capi.siggen(capi.GEN_MULTI_TONE, 8, 6, 1024)
    assert not np.array_equal(a, d)


def test_siggen_tone_rules():
    n = 4096
    x, f, a = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0001, 16, n, with_tones=True)
    assert np.all((f[:, :3] >= -0.5) & (f[:, :3] < 0.5)) and np.all(f[:, 3:] == 0)
    assert np.all((a[:, :3] >= 0.5) & (a[:, :3] <= 2.0))
    # noise power 10 dB below the weakest tone: residual after removing tones
    t = np.arange(n)
    for ch in range(4):
        X = np.fft.fft(x[ch])
        pk = np.sort(np.abs(X))[-3:]
        assert pk.min() > 0.3 * n * a[ch, :3].min()
    x, f, a = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0002, 64, 1024, with_tones=True)
    nt = (a > 0).sum(axis=1)
    assert nt.min() >= 1 and nt.max() <= 8
    assert np.all(a[a > 0] >= 1.0) and np.all(a[a > 0] <= 10.0 + 1e-9)
    x, f, a = capi.siggen(capi.GEN_FRACTIONAL, 0x5EED0005, 32, 8192, with_tones=True)
    for ch in range(32):
        k = np.sort(np.floor(f[ch][a[ch] > 0] * 8192))
        if len(k) > 1:
            assert np.min(np.diff(k)) >= 2  # >= 3 bins apart at integer part (fraction may shift by 1)
    x, f, a = capi.siggen(capi.GEN_DYNAMIC, 0x5EED0004, 8, 16384, with_tones=True)
    assert np.allclose(a[:, :2], [1.0, 1e-3])
    dist = np.abs(f[:, 0] - f[:, 1])
    dist = np.minimum(dist, 1 - dist) * 16384
    assert np.all(dist >= 64)
