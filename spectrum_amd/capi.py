"""ctypes binding of the C ABI in include/spectrum_mi355x.h (libspectrum_mi355x.so).

This is the thinnest host layer: it loads the HIP library, declares every
exported entry point, maps status codes to the reference's exception types
(std::invalid_argument -> ValueError, std::runtime_error -> RuntimeError,
src/fft_func/src/fft_processor_rocm.cpp:311-314,359-361) and offers a small
device-buffer helper so callers need neither torch nor hip-python.

There is no CPU fallback: if the library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspectrum_mi355x.so")

SPX_OK, SPX_EINVAL, SPX_ENOMEM, SPX_EHIP, SPX_ECAP, SPX_ENOTSUP = 0, 1, 2, 3, 4, 5
WINDOW = {"none": 0, "hann": 1, "hamming": 2, "blackman": 3}
DTYPE_C64, DTYPE_C32F16, DTYPE_C32BF16 = 0, 1, 2
PEAKS_THRESHOLD, PEAKS_ALL_MAXIMA = 1, 2
REFINE_NONE, REFINE_PARABOLIC, REFINE_LEARNED = 0, 1, 2
GEN_THREE_TONES, GEN_MULTI_TONE, GEN_DYNAMIC, GEN_FRACTIONAL = 1, 2, 3, 4
COMPUTE_F16, COMPUTE_BF16 = 0, 1

# 32-byte peak record, same layout as the reference MaxValue
MAXVALUE_DTYPE = np.dtype([
    ("index", "<u4"), ("real", "<f4"), ("imag", "<f4"), ("magnitude", "<f4"),
    ("phase", "<f4"), ("freq_offset", "<f4"), ("refined_frequency", "<f4"), ("pad", "<u4"),
])
assert MAXVALUE_DTYPE.itemsize == 32


class SpxFusedDesc(C.Structure):
    _fields_ = [
        ("input", C.c_void_p), ("in_scale", C.c_void_p),
        ("batch", C.c_uint32), ("n_point", C.c_uint32), ("nfft", C.c_uint32),
        ("window", C.c_uint32), ("dtype", C.c_uint32), ("sample_rate", C.c_float),
        ("peak_mode", C.c_uint32), ("nsigma", C.c_float), ("max_peaks", C.c_uint32),
        ("search_start", C.c_uint32), ("search_end", C.c_uint32), ("refine", C.c_uint32),
        ("peaks_out", C.c_void_p), ("counts_out", C.c_void_p), ("gmax_out", C.c_void_p),
        ("stats_out", C.c_void_p),
    ]


class SpxBatchDesc(C.Structure):
    _fields_ = [
        ("input", C.c_void_p), ("offsets", C.c_void_p), ("lengths", C.c_void_p),
        ("n_chan", C.c_uint32), ("window", C.c_uint32), ("dtype", C.c_uint32),
        ("sample_rate", C.c_float), ("nsigma", C.c_float), ("max_peaks", C.c_uint32), ("refine", C.c_uint32),
        ("peaks_out", C.c_void_p), ("counts_out", C.c_void_p), ("gmax_out", C.c_void_p),
        ("stats_out", C.c_void_p), ("tune_cache", C.c_char_p), ("flags", C.c_uint32),
    ]


BATCH_NO_GRAPH = 1

# (name, restype, argtypes) for every symbol declared in include/spectrum_mi355x.h
_P, _U32, _U64, _I, _F = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float
SIGNATURES = [
    ("spx_create", _I, [_I, C.POINTER(C.c_void_p)]),
    ("spx_destroy", _I, [_P]),
    ("spx_set_stream", _I, [_P, _P]),
    ("spx_get_stream", _P, [_P]),
    ("spx_synchronize", _I, [_P]),
    ("spx_last_error", C.c_char_p, [_P]),
    ("spx_device_info", _I, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    ("spx_alloc", _I, [_P, _U64, C.POINTER(C.c_void_p)]),
    ("spx_free", _I, [_P, _P]),
    ("spx_memcpy_h2d", _I, [_P, _P, _P, _U64]),
    ("spx_memcpy_d2h", _I, [_P, _P, _P, _U64]),
    ("spx_fft_c2c", _I, [_P, _P, _P, _U32, _U32, _U32, _U32]),
    ("spx_fft_mag_phase", _I, [_P, _P, _P, _U32, _U32, _U32, _U32]),
    ("spx_fft_magnitude", _I, [_P, _P, _P, _U32, _U32, _U32, _U32, _I, _F]),
    ("spx_mag_phase", _I, [_P, _P, _P, _U64]),
    ("spx_magnitude", _I, [_P, _P, _P, _U64, _F, _I]),
    ("spx_peaks_top", _I, [_P, _P, _P, _U32, _U32, _U32, _F, _I]),
    ("spx_fused_peaks_top", _I, [_P, _P, _P, _U32, _U32, _U32, _U32, _U32, _F, _I]),
    ("spx_all_maxima", _I, [_P, _P, _P, _P, _U32, _U32, _F, _U32, _U32, _U32]),
    ("spx_fused_peaks", _I, [_P, C.POINTER(SpxFusedDesc)]),
    ("spx_fused_peaks_noalloc", _I, [_P, C.POINTER(SpxFusedDesc)]),
    ("spx_fft_c2c_host", _I, [_P, _P, _P, _U32, _U32, _U32, _U32]),
    ("spx_fft_mag_phase_host", _I, [_P, _P, _P, _U32, _U32, _U32, _U32]),
    ("spx_fft_magnitude_host", _I, [_P, _P, _P, _U32, _U32, _U32, _U32, _I, _F]),
    ("spx_fused_peaks_host", _I, [_P, C.POINTER(SpxFusedDesc)]),
    ("spx_set_host_chunk", _I, [_P, _U64]),
    ("spx_host_alloc", _I, [_P, _U64, C.POINTER(C.c_void_p)]),
    ("spx_host_free", _I, [_P, _P]),
    ("spx_mlp_param_count", _I, []),
    ("spx_mlp_deployed_size", _I, []),
    ("spx_mlp_init", _I, [_P, _P, _U64]),
    ("spx_mlp_grad", _I, [_P, _P, _P, _P, _U32, _P]),
    ("spx_mlp_train", _I, [_P, _P, _P, _P, _P, _P, _P, _U32, _U32, _U32, _U32, _F, _U64, _P]),
    ("spx_mlp_export", _I, [_P, _P, _P]),
    ("spx_mlp_predict", _I, [_P, _P, _P, _U32, _P]),
    ("spx_set_refiner", _I, [_P, _P]),
    ("spx_mlp_dataset", _I, [_P, _P, _P, _P, _P, _U32, _U32, _U32, _P, _P, _P, _P]),
    ("spx_fft_mixed", _I, [_P, _P, _P, _U32, _U32, _U32, _U32, _U32]),
    ("spx_fused_peaks_mixed", _I, [_P, C.POINTER(SpxFusedDesc), _U32]),
    ("spx_batch_plan_create", _I, [_P, C.POINTER(SpxBatchDesc), C.POINTER(C.c_void_p)]),
    ("spx_batch_execute", _I, [_P]),
    ("spx_batch_plan_destroy", _I, [_P]),
    ("spx_batch_plan_info", _I, [_P, C.c_char_p, C.c_size_t]),
    ("spx_batch_autotune", _I, [_P, C.POINTER(SpxBatchDesc), C.c_char_p]),
    ("spx_siggen", _I, [_U32, _U64, _U32, _U32, _U32, _P, _U32, _P, _P, _I]),
    ("spx_siggen_lengths", _I, [_U64, _U32, _U32, _U32, _P]),
    ("spx_siggen_uneven", _I, [_U32, _U64, _U32, _P, _P, _P, _I]),
    ("spx_version", C.c_char_p, []),
]

_lib = None


def load(path: Optional[str] = None) -> C.CDLL:
    """Load libspectrum_mi355x.so and declare its signatures (raises if absent)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def exported_symbols():
    return [s[0] for s in SIGNATURES]


def _ptr(x) -> Optional[int]:
    """Device address of a DeviceBuffer, a torch tensor, an int or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, DeviceBuffer):
        return x.ptr
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    raise TypeError(f"cannot take a device address of {type(x)!r}")


class SpxError(RuntimeError):
    pass


class Context:
    """One spx_ctx on one device (replaces the reference's IBackend + GpuContext)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.spx_create(device, C.byref(h))
        if rc != SPX_OK:
            raise RuntimeError(f"spx_create(device={device}) failed with status {rc} (no usable HIP device?)")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.spx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- status handling
    def check(self, rc: int, what: str):
        if rc == SPX_OK:
            return
        msg = (self.lib.spx_last_error(self.h) or b"").decode(errors="replace")
        text = f"{what}: {msg} (status {rc})"
        if rc == SPX_EINVAL:
            raise ValueError(text)
        raise SpxError(text)

    def info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int()
        hbm = C.c_uint64()
        self.check(self.lib.spx_device_info(self.h, name, 256, C.byref(cu), C.byref(hbm)), "spx_device_info")
        return {"name": name.value.decode(), "cu_count": cu.value, "hbm_bytes": hbm.value}

    def set_stream(self, stream: Optional[int]):
        self.check(self.lib.spx_set_stream(self.h, stream), "spx_set_stream")

    def stream(self) -> int:
        return int(self.lib.spx_get_stream(self.h) or 0)

    def sync(self):
        self.check(self.lib.spx_synchronize(self.h), "spx_synchronize")

    # -- memory
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, arr.nbytes)
        b.upload(arr)
        return b

    # -- entry points (asynchronous; call sync() before reading results)
    def fft_c2c(self, inp, out, batch, n_point, nfft, window=0):
        self.check(self.lib.spx_fft_c2c(self.h, _ptr(inp), _ptr(out), batch, n_point, nfft, window), "spx_fft_c2c")

    def fft_mag_phase(self, inp, out, batch, n_point, nfft, window=0):
        self.check(self.lib.spx_fft_mag_phase(self.h, _ptr(inp), _ptr(out), batch, n_point, nfft, window),
                   "spx_fft_mag_phase")

    def fft_magnitude(self, inp, out, batch, n_point, nfft, window=0, squared=False, scale=1.0):
        self.check(self.lib.spx_fft_magnitude(self.h, _ptr(inp), _ptr(out), batch, n_point, nfft, window,
                                              int(bool(squared)), float(scale)), "spx_fft_magnitude")

    def mag_phase(self, inp, out, total):
        self.check(self.lib.spx_mag_phase(self.h, _ptr(inp), _ptr(out), total), "spx_mag_phase")

    def magnitude(self, inp, out, total, inv_n=1.0, squared=False):
        self.check(self.lib.spx_magnitude(self.h, _ptr(inp), _ptr(out), total, float(inv_n), int(bool(squared))),
                   "spx_magnitude")

    def peaks_top(self, spec, out, batch, nfft, search_range, fs, two_peaks=False):
        self.check(self.lib.spx_peaks_top(self.h, _ptr(spec), _ptr(out), batch, nfft, search_range, float(fs),
                                          int(bool(two_peaks))), "spx_peaks_top")

    def fused_peaks_top(self, inp, out, batch, n_point, nfft, search_range, fs, two_peaks=False, window=0):
        """Pad + window + FFT + reference ONE_PEAK (4 records per channel) or
        TWO_PEAKS (8) in the transform's launch; `out` holds batch * 4 (8)
        MAXVALUE_DTYPE records."""
        self.check(self.lib.spx_fused_peaks_top(self.h, _ptr(inp), _ptr(out), batch, n_point, nfft, window,
                                                search_range, float(fs), int(bool(two_peaks))),
                   "spx_fused_peaks_top")

    def all_maxima(self, spec, out, counts, batch, nfft, fs, s0, s1, cap):
        self.check(self.lib.spx_all_maxima(self.h, _ptr(spec), _ptr(out), _ptr(counts), batch, nfft, float(fs),
                                           s0, s1, cap), "spx_all_maxima")

    def fused_peaks(self, inp, batch, n_point, nfft, peaks, counts, *, window=1, dtype=DTYPE_C64, fs=1.0,
                    peak_mode=PEAKS_THRESHOLD, nsigma=6.0, max_peaks=32, s0=0, s1=0, refine=REFINE_PARABOLIC,
                    gmax=None, stats=None, in_scale=None, noalloc=False):
        """spx_fused_peaks, or with noalloc=True spx_fused_peaks_noalloc (fixed
        workspace, no device allocation; the context must hold this length's
        tables already)."""
        d = SpxFusedDesc()
        d.input = _ptr(inp)
        d.in_scale = _ptr(in_scale)
        d.batch, d.n_point, d.nfft = batch, n_point, nfft
        d.window, d.dtype, d.sample_rate = window, dtype, float(fs)
        d.peak_mode, d.nsigma, d.max_peaks = peak_mode, float(nsigma), max_peaks
        d.search_start, d.search_end, d.refine = s0, s1, refine
        d.peaks_out, d.counts_out = _ptr(peaks), _ptr(counts)
        d.gmax_out, d.stats_out = _ptr(gmax), _ptr(stats)
        if noalloc:
            self.check(self.lib.spx_fused_peaks_noalloc(self.h, C.byref(d)), "spx_fused_peaks_noalloc")
        else:
            self.check(self.lib.spx_fused_peaks(self.h, C.byref(d)), "spx_fused_peaks")

    # -- host-resident data (numpy in, numpy out; synchronous, streamed in chunks)
    def set_host_chunk(self, nbytes: int):
        self.check(self.lib.spx_set_host_chunk(self.h, int(nbytes)), "spx_set_host_chunk")

    def pinned(self, shape, dtype) -> "PinnedArray":
        """A numpy array in pinned host memory (no staging copy in the *_host calls)."""
        return PinnedArray(self, shape, dtype)

    @staticmethod
    def _host_in(x: np.ndarray, dtype) -> np.ndarray:
        x = np.ascontiguousarray(x)
        if x.dtype != dtype:
            raise TypeError(f"expected {np.dtype(dtype)} input, got {x.dtype}")
        return x

    def fft_c2c_host(self, x: np.ndarray, nfft: int, window: int = 0, out: Optional[np.ndarray] = None):
        """[batch][n_point] complex64 host array -> [batch][nfft] complex64 spectra."""
        x = self._host_in(x, np.complex64)
        B, n = x.shape
        out = np.empty((B, nfft), np.complex64) if out is None else out
        self.check(self.lib.spx_fft_c2c_host(self.h, x.ctypes.data, out.ctypes.data, B, n, nfft, window),
                   "spx_fft_c2c_host")
        return out

    def fft_mag_phase_host(self, x: np.ndarray, nfft: int, window: int = 0):
        """-> [batch][nfft][2] float32 {|X|, arg X}."""
        x = self._host_in(x, np.complex64)
        B, n = x.shape
        out = np.empty((B, nfft, 2), np.float32)
        self.check(self.lib.spx_fft_mag_phase_host(self.h, x.ctypes.data, out.ctypes.data, B, n, nfft, window),
                   "spx_fft_mag_phase_host")
        return out

    def fft_magnitude_host(self, x: np.ndarray, nfft: int, window: int = 0, squared=False, scale=1.0):
        x = self._host_in(x, np.complex64)
        B, n = x.shape
        out = np.empty((B, nfft), np.float32)
        self.check(self.lib.spx_fft_magnitude_host(self.h, x.ctypes.data, out.ctypes.data, B, n, nfft, window,
                                                   int(bool(squared)), float(scale)), "spx_fft_magnitude_host")
        return out

    def fused_peaks_host(self, x: np.ndarray, nfft: int, *, window=1, fs=1.0, peak_mode=PEAKS_THRESHOLD,
                         nsigma=6.0, max_peaks=32, s0=0, s1=0, refine=REFINE_PARABOLIC, with_gmax=True,
                         with_stats=True, dtype=DTYPE_C64, in_scale: Optional[np.ndarray] = None):
        """Host signal [batch][n_point] (complex64, or [batch][n_point][2] float16 / uint16-bf16 for the
        fp16 / bf16 dtypes) -> (records [batch][max_peaks], counts [batch][2], gmax [batch] or None,
        stats [batch][4] or None), all numpy."""
        x = np.ascontiguousarray(x)
        B, n = x.shape[0], x.shape[1]
        pk = np.zeros((B, max_peaks), MAXVALUE_DTYPE)
        ct = np.zeros((B, 2), np.uint32)
        gm = np.zeros(B, MAXVALUE_DTYPE) if with_gmax else None
        st = np.zeros((B, 4), np.float32) if with_stats else None
        sc = None if in_scale is None else np.ascontiguousarray(in_scale, dtype=np.float32)
        d = SpxFusedDesc()
        d.input = x.ctypes.data
        d.in_scale = None if sc is None else sc.ctypes.data
        d.batch, d.n_point, d.nfft = B, n, nfft
        d.window, d.dtype, d.sample_rate = window, dtype, float(fs)
        d.peak_mode, d.nsigma, d.max_peaks = peak_mode, float(nsigma), max_peaks
        d.search_start, d.search_end, d.refine = s0, s1, refine
        d.peaks_out, d.counts_out = pk.ctypes.data, ct.ctypes.data
        d.gmax_out = None if gm is None else gm.ctypes.data
        d.stats_out = None if st is None else st.ctypes.data
        self.check(self.lib.spx_fused_peaks_host(self.h, C.byref(d)), "spx_fused_peaks_host")
        return pk, ct, gm, st

    # -- learned sub-bin refiner
    def set_refiner(self, deployed):
        self.check(self.lib.spx_set_refiner(self.h, _ptr(deployed)), "spx_set_refiner")

    def mlp_init(self, params, seed):
        self.check(self.lib.spx_mlp_init(self.h, _ptr(params), seed), "spx_mlp_init")

    def mlp_grad(self, params, X, Y, n, grad):
        self.check(self.lib.spx_mlp_grad(self.h, _ptr(params), _ptr(X), _ptr(Y), n, _ptr(grad)), "spx_mlp_grad")

    def mlp_train(self, params, m, v, grad, X, Y, n, batch, steps, step0, lr, seed):
        losses = np.zeros(steps, dtype=np.float32)
        self.check(self.lib.spx_mlp_train(self.h, _ptr(params), _ptr(m), _ptr(v), _ptr(grad), _ptr(X), _ptr(Y), n,
                                          batch, steps, step0, float(lr), seed, losses.ctypes.data),
                   "spx_mlp_train")
        return losses

    def mlp_export(self, params, deployed):
        self.check(self.lib.spx_mlp_export(self.h, _ptr(params), _ptr(deployed)), "spx_mlp_export")

    def mlp_predict(self, deployed, X, n, out):
        self.check(self.lib.spx_mlp_predict(self.h, _ptr(deployed), _ptr(X), n, _ptr(out)), "spx_mlp_predict")

    def mlp_dataset(self, spec, mag16, peaks, counts, batch, nfft, max_peaks, feat, label, parab, valid):
        self.check(self.lib.spx_mlp_dataset(self.h, _ptr(spec), _ptr(mag16), _ptr(peaks), _ptr(counts), batch, nfft,
                                            max_peaks, _ptr(feat), _ptr(label), _ptr(parab), _ptr(valid)),
                   "spx_mlp_dataset")

    # -- mixed precision (fp16 / bf16 storage, MFMA DFT-as-GEMM), N = 16384
    def fft_mixed(self, inp, out, batch, nfft=16384, window=1, dtype=DTYPE_C32F16, compute=COMPUTE_F16):
        self.check(self.lib.spx_fft_mixed(self.h, _ptr(inp), _ptr(out), batch, nfft, window, dtype, compute),
                   "spx_fft_mixed")

    def fused_peaks_mixed(self, inp, batch, nfft, peaks, counts, *, compute=COMPUTE_F16, window=1,
                          dtype=DTYPE_C32F16, fs=1.0, nsigma=6.0, max_peaks=32, s0=0, s1=0,
                          refine=REFINE_PARABOLIC, gmax=None, stats=None, in_scale=None):
        d = SpxFusedDesc()
        d.input = _ptr(inp)
        d.in_scale = _ptr(in_scale)
        d.batch, d.n_point, d.nfft = batch, nfft, nfft
        d.window, d.dtype, d.sample_rate = window, dtype, float(fs)
        d.peak_mode, d.nsigma, d.max_peaks = PEAKS_THRESHOLD, float(nsigma), max_peaks
        d.search_start, d.search_end, d.refine = s0, s1, refine
        d.peaks_out, d.counts_out = _ptr(peaks), _ptr(counts)
        d.gmax_out, d.stats_out = _ptr(gmax), _ptr(stats)
        self.check(self.lib.spx_fused_peaks_mixed(self.h, C.byref(d), compute), "spx_fused_peaks_mixed")

    # -- uneven-length batch
    def batch_desc(self, inp, offsets: np.ndarray, lengths: np.ndarray, peaks, counts, *, window=1,
                   dtype=DTYPE_C64, fs=1.0, nsigma=6.0, max_peaks=64, refine=REFINE_PARABOLIC, gmax=None,
                   stats=None, tune_cache: Optional[str] = None, graph=True):
        """Descriptor for spx_batch_*; keeps the host arrays alive on the returned object."""
        d = SpxBatchDesc()
        d._offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        d._lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
        d.input = _ptr(inp)
        d.offsets = d._offsets.ctypes.data
        d.lengths = d._lengths.ctypes.data
        d.n_chan = len(d._lengths)
        d.window, d.dtype, d.sample_rate = window, dtype, float(fs)
        d.nsigma, d.max_peaks, d.refine = float(nsigma), max_peaks, refine
        d.peaks_out, d.counts_out = _ptr(peaks), _ptr(counts)
        d.gmax_out, d.stats_out = _ptr(gmax), _ptr(stats)
        d._cache = tune_cache.encode() if tune_cache else None
        d.tune_cache = d._cache
        d.flags = 0 if graph else BATCH_NO_GRAPH
        return d

    def batch_plan(self, desc: SpxBatchDesc) -> "BatchPlan":
        return BatchPlan(self, desc)

    def batch_autotune(self, desc: SpxBatchDesc, cache_path: str):
        self.check(self.lib.spx_batch_autotune(self.h, C.byref(desc), cache_path.encode()), "spx_batch_autotune")


class BatchPlan:
    """spx_batch_plan: work list + scratch + captured graph for one uneven batch."""

    def __init__(self, ctx: Context, desc: SpxBatchDesc):
        self.ctx = ctx
        self.desc = desc
        h = C.c_void_p()
        ctx.check(ctx.lib.spx_batch_plan_create(ctx.h, C.byref(desc), C.byref(h)), "spx_batch_plan_create")
        self.h = h

    def execute(self):
        self.ctx.check(self.ctx.lib.spx_batch_execute(self.h), "spx_batch_execute")

    def info(self) -> dict:
        import json

        buf = C.create_string_buffer(8192)
        self.ctx.check(self.ctx.lib.spx_batch_plan_info(self.h, buf, 8192), "spx_batch_plan_info")
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.spx_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """Pinned host memory from spx_host_alloc, viewed as a numpy array (.array);
    freed with the object (keep it alive while .array is in use)."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        p = C.c_void_p()
        ctx.check(ctx.lib.spx_host_alloc(ctx.h, nbytes, C.byref(p)), "spx_host_alloc")
        self.ptr = p.value
        buf = (C.c_char * max(nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)

    def free(self):
        if getattr(self, "ptr", None):
            self.ctx.lib.spx_host_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBuffer:
    """Caller-owned device allocation made through spx_alloc (freed on close/GC)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx.check(ctx.lib.spx_alloc(ctx.h, self.nbytes, C.byref(p)), "spx_alloc")
        self.ptr = int(p.value or 0)

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        self.ctx.check(self.ctx.lib.spx_memcpy_h2d(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes), "h2d")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested more bytes than the buffer holds")
        self.ctx.check(self.ctx.lib.spx_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx.lib.spx_free(self.ctx.h, self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def siggen(kind: int, seed: int, n_chan: int, n_point: int, chan_offset: int = 0, max_tones: int = 16,
           threads: int = 0, with_tones: bool = False):
    """Host-side deterministic generator (spx_siggen); returns complex64 [n_chan, n_point]."""
    lib = load()
    out = np.empty((n_chan, n_point), dtype=np.complex64)
    tf = np.zeros((n_chan, max_tones), dtype=np.float64) if with_tones else None
    ta = np.zeros((n_chan, max_tones), dtype=np.float64) if with_tones else None
    rc = lib.spx_siggen(kind, seed, n_chan, n_point, chan_offset, out.ctypes.data, max_tones,
                        tf.ctypes.data if tf is not None else None, ta.ctypes.data if ta is not None else None,
                        threads)
    if rc != SPX_OK:
        raise ValueError(f"spx_siggen failed with status {rc}")
    return (out, tf, ta) if with_tones else out


def siggen_lengths(seed: int, n_chan: int, lo_log2: int = 10, hi_log2: int = 17) -> np.ndarray:
    """Per-channel power-of-two lengths 2^(lo + U{0..hi-lo}) (spx_siggen_lengths)."""
    out = np.empty(n_chan, dtype=np.uint32)
    rc = load().spx_siggen_lengths(seed, n_chan, lo_log2, hi_log2, out.ctypes.data)
    if rc != SPX_OK:
        raise ValueError(f"spx_siggen_lengths failed with status {rc}")
    return out


def length_major_offsets(lengths: np.ndarray):
    """Offsets that pack channels grouped by length (ascending length, then channel id).

    Returns (offsets[n_chan] in complex elements, total elements, {length: (first offset, channel ids)})."""
    lengths = np.asarray(lengths, dtype=np.uint64)
    order = np.lexsort((np.arange(len(lengths)), lengths))
    offsets = np.empty(len(lengths), dtype=np.uint64)
    pos = 0
    groups = {}
    for c in order:
        L = int(lengths[c])
        if L not in groups:
            groups[L] = (pos, [])
        groups[L][1].append(int(c))
        offsets[c] = pos
        pos += L
    return offsets, pos, {L: (o, np.array(ids, dtype=np.int64)) for L, (o, ids) in groups.items()}


def siggen_uneven(kind: int, seed: int, lengths: np.ndarray, offsets: np.ndarray, total: int, threads: int = 0):
    """Channels of individual lengths packed at `offsets` (spx_siggen_uneven); complex64 [total]."""
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    out = np.empty(int(total), dtype=np.complex64)
    rc = load().spx_siggen_uneven(kind, seed, len(lengths), lengths.ctypes.data, offsets.ctypes.data,
                                  out.ctypes.data, threads)
    if rc != SPX_OK:
        raise ValueError(f"spx_siggen_uneven failed with status {rc}")
    return out


def to_half_storage(x: np.ndarray, dtype: int) -> np.ndarray:
    """complex64 -> packed 16-bit complex storage (uint32 per element: re in the low half, im in the high half),
    IEEE fp16 (DTYPE_C32F16) or bfloat16 (DTYPE_C32BF16), round to nearest even."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    f = x.view(np.float32).reshape(x.shape + (2,))
    if dtype == DTYPE_C32F16:
        h = f.astype(np.float16).view(np.uint16).astype(np.uint32)
    elif dtype == DTYPE_C32BF16:
        u = f.view(np.uint32).astype(np.uint64)
        h = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) & 0xFFFF
    else:
        raise ValueError("dtype must be DTYPE_C32F16 or DTYPE_C32BF16")
    return (h[..., 0] | (h[..., 1] << 16)).astype(np.uint32)


def from_half_storage(u: np.ndarray, dtype: int) -> np.ndarray:
    """Inverse of to_half_storage (exact): packed 16-bit complex -> complex64."""
    u = np.asarray(u, dtype=np.uint32)
    lo, hi = (u & 0xFFFF).astype(np.uint16), (u >> 16).astype(np.uint16)
    if dtype == DTYPE_C32F16:
        re, im = lo.view(np.float16).astype(np.float32), hi.view(np.float16).astype(np.float32)
    else:
        re = (lo.astype(np.uint32) << 16).view(np.float32)
        im = (hi.astype(np.uint32) << 16).view(np.float32)
    return (re + 1j * im).astype(np.complex64)
