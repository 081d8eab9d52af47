"""Python side of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's cpu_baseline leg may import this
module, and only as the checker.  It wraps:
  * _build/liboracle.so  -- the plain-C FP64 restatement (spx_oracle.c)
  * _ref/libref_cpufft.so, _ref/libref_pocketfft_mt.so -- the reference's own
    CPU FFT (include/spectrum/utils/cpu_fft.hpp over third_party/pocketfft),
    compiled from the reference tree by oracle/Makefile (present only where the
    reference was available at build time; the built files travel with the tree)
and provides the tie-tolerant peak-list comparator.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
REF_ST = os.path.join(HERE, "_ref", "libref_cpufft.so")
REF_MT = os.path.join(HERE, "_ref", "libref_pocketfft_mt.so")

MAXVALUE_DTYPE = np.dtype([
    ("index", "<u4"), ("real", "<f4"), ("imag", "<f4"), ("magnitude", "<f4"),
    ("phase", "<f4"), ("freq_offset", "<f4"), ("refined_frequency", "<f4"), ("pad", "<u4"),
])

_lib = None
_ref_st = None
_ref_mt = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle lib`")
        L = C.CDLL(LIB)
        P, I, D, F, U = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_uint32
        L.orc_fft.argtypes = [P, P, I]
        L.orc_dft_brute.argtypes = [P, P, I]
        L.orc_fft_batch.argtypes = [P, P, I, I, I, P]
        L.orc_window.argtypes = [I, I, P]
        L.orc_window.restype = None
        L.orc_fused_peaks.argtypes = [P, I, D, I, I, I, D, I, I, P, P, P, P, P, P]
        L.orc_peaks_top.argtypes = [P, U, U, F, I, P]
        L.orc_all_maxima.argtypes = [P, U, F, U, U, U, P, P]
        _lib = L
    return _lib


def have_reference() -> bool:
    return os.path.exists(REF_ST) and os.path.exists(REF_MT)


def _ref():
    global _ref_st, _ref_mt
    if _ref_st is None:
        if not have_reference():
            raise RuntimeError("oracle/_ref not built (needs the reference tree at build time)")
        _ref_st = C.CDLL(REF_ST)
        _ref_st.ref_cpufft_c2c.argtypes = [C.c_void_p, C.c_void_p, C.c_ulong]
        _ref_st.ref_cpufft_c2c_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_ulong, C.c_ulong]
        _ref_mt = C.CDLL(REF_MT)
        _ref_mt.ref_pocketfft_c2c_mt.argtypes = [C.c_void_p, C.c_void_p, C.c_ulong, C.c_ulong, C.c_ulong]
    return _ref_st, _ref_mt


# ---------------------------------------------------------------- transforms
def fft64(x: np.ndarray) -> np.ndarray:
    """FP64 forward FFT of the last axis (power-of-two length), unnormalised."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    n = x.shape[-1]
    flat_in, flat_out = x.reshape(-1, n), out.reshape(-1, n)
    for i in range(flat_in.shape[0]):
        rc = lib().orc_fft(flat_in[i].ctypes.data, flat_out[i].ctypes.data, n)
        if rc:
            raise ValueError("orc_fft failed")
    return out


def dft_brute(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_dft_brute(x.ctypes.data, out.ctypes.data, x.shape[-1])
    return out


def window(kind: int, n: int) -> np.ndarray:
    w = np.empty(n, dtype=np.float64)
    lib().orc_window(kind, n, w.ctypes.data)
    return w


def spectrum64(x: np.ndarray, nfft: Optional[int] = None, win: int = 0) -> np.ndarray:
    """Window (symmetric, over n_point) + zero-pad + FP64 FFT of [batch, n_point] complex64."""
    x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex64)
    b, n = x.shape
    nfft = nfft or n
    out = np.empty((b, nfft), dtype=np.complex128)
    w = window(win, n) if win else None
    rc = lib().orc_fft_batch(x.ctypes.data, out.ctypes.data, b, n, nfft, w.ctypes.data if w is not None else None)
    if rc:
        raise ValueError("orc_fft_batch failed")
    return out


def ref_cpufft(x: np.ndarray) -> np.ndarray:
    """The reference's CpuFFT::ForwardC2C (pocketfft float, single thread) per row."""
    st, _ = _ref()
    x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex64)
    out = np.empty_like(x)
    if st.ref_cpufft_c2c_batch(x.ctypes.data, out.ctypes.data, x.shape[0], x.shape[1]):
        raise RuntimeError("ref_cpufft_c2c failed")
    return out


def ref_pocketfft_mt(x: np.ndarray, nthreads: int) -> np.ndarray:
    _, mt = _ref()
    x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex64)
    out = np.empty_like(x)
    if mt.ref_pocketfft_c2c_mt(x.ctypes.data, out.ctypes.data, x.shape[0], x.shape[1], nthreads):
        raise RuntimeError("ref_pocketfft_c2c_mt failed")
    return out


# ---------------------------------------------------------------- peak rules
def fused_peaks(X: np.ndarray, fs: float, *, mode: int = 1, nsigma: float = 6.0, kcap: int = 32, s0: int = 0,
                s1: int = 0, refine: int = 1):
    """Threshold (mode 1) or AllMaxima (mode 2) rules on one complex spectrum.

    Returns (records[stored], found, gmax_record, stats[4], mags[nfft])."""
    Xf = np.ascontiguousarray(X, dtype=np.complex64)
    n = Xf.shape[-1]
    if mode == 2:
        s0 = s0 or 1
        s1 = s1 or n // 2
        s1 = min(s1, n - 1)
    elif s0 == 0 and s1 == 0:
        s1 = n
    recs = np.zeros(max(kcap, 1), dtype=MAXVALUE_DTYPE)
    gm = np.zeros(1, dtype=MAXVALUE_DTYPE)
    stored, found = C.c_int(), C.c_int()
    stats = np.zeros(4, dtype=np.float64)
    mags = np.zeros(n, dtype=np.float64)
    rc = lib().orc_fused_peaks(Xf.ctypes.data, n, float(fs), s0, s1, mode, float(nsigma), kcap, refine,
                               recs.ctypes.data, C.byref(stored), C.byref(found), gm.ctypes.data, stats.ctypes.data,
                               mags.ctypes.data)
    if rc:
        raise ValueError("orc_fused_peaks failed")
    return recs[: stored.value], found.value, gm[0], stats, mags


def peaks_top(X: np.ndarray, search_range: int, fs: float, two_peaks: bool = False) -> np.ndarray:
    Xf = np.ascontiguousarray(X, dtype=np.complex64)
    out = np.zeros(8 if two_peaks else 4, dtype=MAXVALUE_DTYPE)
    lib().orc_peaks_top(Xf.ctypes.data, Xf.shape[-1], search_range, float(fs), int(two_peaks), out.ctypes.data)
    return out


def all_maxima(X: np.ndarray, fs: float, s0: int, s1: int, cap: int):
    Xf = np.ascontiguousarray(X, dtype=np.complex64)
    out = np.zeros(max(cap, 1), dtype=MAXVALUE_DTYPE)
    cnt = C.c_uint32()
    lib().orc_all_maxima(Xf.ctypes.data, Xf.shape[-1], float(fs), s0, s1, cap, out.ctypes.data, C.byref(cnt))
    return out[: min(cnt.value, cap)], cnt.value


# ---------------------------------------------------------------- comparator
def excused_bins(mags: np.ndarray, thr: float, tau: float, nfft: int = 0, s0: int = 0, s1: int = 0, circular=True):
    """Bins whose peak decision is within tau of a tie: a local-maximum test
    (|m[k] - m[k-1]| or |m[k] - m[k+1]| <= tau, circular neighbours) or the
    threshold test (|m[k] - thr| <= tau).  Range decisions are exact."""
    m = np.asarray(mags, dtype=np.float64)
    left = np.roll(m, 1)
    right = np.roll(m, -1)
    close = (np.abs(m - left) <= tau) | (np.abs(m - right) <= tau)
    if np.isfinite(thr):
        close |= np.abs(m - thr) <= tau
    return close


def possible_survivors(mags: np.ndarray, thr: float, tau: float) -> np.ndarray:
    """Bins that are a peak (strict local maximum above thr) under some
    rounding of the spectrum within tau: m[k] + tau > m[k-1] - tau,
    m[k] + tau > m[k+1] - tau and m[k] + tau > thr."""
    m = np.asarray(mags, dtype=np.float64)
    ok = (m + tau > np.roll(m, 1) - tau) & (m + tau > np.roll(m, -1) - tau)
    if np.isfinite(thr):
        ok &= m + tau > thr
    return ok


def compare_peak_lists(idx_gpu, idx_ref, mags, thr, tau, kcap, found_gpu=None, found_ref=None, s0=0, s1=0):
    """Tie-tolerant comparison of two ascending peak-index lists, each the
    first kcap peaks of its spectrum (a list shorter than kcap holds all).

    A bin is excused when its decision is within tau of a tie (excused_bins).
    Let L be the smaller last bin of the lists that are full (kcap entries);
    up to L both lists are complete, so their non-excused bins must be equal
    there.  Beyond L only the full list has entries and nothing is compared.
    found_gpu / found_ref (uncapped counts, optional) may differ by at most
    the number of excused bins in [s0, s1) (or the whole spectrum) that are
    possible survivors.  Returns (ok, detail)."""
    a = [int(i) for i in idx_gpu]
    b = [int(i) for i in idx_ref]
    m = np.asarray(mags, dtype=np.float64)
    exc = excused_bins(m, thr, tau)
    if found_gpu is not None and found_ref is not None and int(found_gpu) != int(found_ref):
        lo, hi = (s0, s1) if s1 > s0 else (0, len(m))
        slack = int(np.count_nonzero(exc[lo:hi] & possible_survivors(m, thr, tau)[lo:hi]))
        if abs(int(found_gpu) - int(found_ref)) > slack:
            return False, f"found {found_gpu} vs {found_ref}, more than the {slack} excused possible survivors"
    if a == b:
        return True, "equal"
    if a != sorted(set(a)):
        return False, "gpu list not strictly ascending"
    la = a[-1] if len(a) >= kcap and kcap > 0 else np.inf
    lb = b[-1] if len(b) >= kcap and kcap > 0 else np.inf
    L = min(la, lb)
    ca = [k for k in a if k <= L and not exc[k]]
    cb = [k for k in b if k <= L and not exc[k]]
    if ca != cb:
        bad = sorted(set(ca) ^ set(cb))
        return False, f"unexcused {bad[:10]} up to bin {L} (gpu {len(a)}, ref {len(b)}, kcap {kcap})"
    diff = sorted(set(a) ^ set(b))
    return True, f"excused {[k for k in diff if exc[k]][:10]}" + (f", tail past bin {L}" if np.isfinite(L) else "")
