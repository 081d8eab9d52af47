"""The pybind11 module `dsp_spectrum` on the CPU: its host FFT utilities
(python/py_cpu_fft.hpp:28-106 of the reference, tested there by
python/t_cpu_fft.py, whose cases are restated here) and the reference's class,
method and keyword names (python/dsp_spectrum_module.cpp:35-68,
py_fft_processor_rocm.hpp:183-225, py_spectrum_maxima_finder_rocm.hpp:296-336,
py_complex_to_mag_rocm.hpp:126-147)."""
import re

import numpy as np
import pytest

from spectrum_amd import dsp_spectrum as ds
from oracle import oracle as orc


def test_c2c_forward_peak():
    N, fs, freq = 1024, 1000.0, 100.0
    sig = np.exp(2j * np.pi * freq * np.arange(N) / fs).astype(np.complex64)
    X = ds.cpu_fft_c2c(sig)
    assert X.shape == (N,) and X.dtype == np.complex64
    assert abs(int(np.argmax(np.abs(X[: N // 2]))) * fs / N - freq) < fs / N


def test_c2c_roundtrip():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(512) + 1j * rng.standard_normal(512)).astype(np.complex64)
    assert np.max(np.abs(ds.cpu_ifft_c2c(ds.cpu_fft_c2c(x)) - x)) < 1e-5


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 1000, 2048, 4097])
def test_c2c_vs_fp64_oracle(n):
    """Any length (power of two and Bluestein paths) against the FP64 brute-force DFT."""
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ref = orc.dft_brute(x)
    got = ds.cpu_fft_c2c(x)
    assert np.max(np.abs(got - ref)) <= 2e-6 * max(1.0, np.abs(ref).max())
    back = ds.cpu_ifft_c2c(ref.astype(np.complex64))
    assert np.max(np.abs(back - x)) <= 1e-5 * max(1.0, np.abs(x).max())


def test_c2c_vs_reference_cpufft():
    """Same vectors through the reference's own CpuFFT (pocketfft, built from its sources)."""
    if not orc.have_reference():
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(7)
    for n in (256, 2048, 65536):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        ref = orc.ref_cpufft(x[None])[0]
        got = ds.cpu_fft_c2c(x)
        assert np.max(np.abs(got - ref)) / np.abs(ref).max() < 1e-5


def test_r2c_half_and_full():
    N, fs, freq = 2048, 4000.0, 500.0
    sig = np.sin(2 * np.pi * freq * np.arange(N) / fs).astype(np.float32)
    half = ds.cpu_fft_r2c(sig)
    assert half.shape == (N // 2 + 1,)
    assert abs(int(np.argmax(np.abs(half))) * fs / N - freq) < fs / N
    full = ds.cpu_fft_r2c_full(sig)
    assert full.shape == (N,)
    k = np.arange(1, N // 2)
    assert np.max(np.abs(full[k] - np.conj(full[N - k]))) < 1e-5
    assert np.allclose(full[: N // 2 + 1], half)
    x = np.random.default_rng(3).standard_normal(4096).astype(np.float32)
    assert np.max(np.abs(ds.cpu_fft_r2c(x) - np.fft.rfft(x.astype(np.float64)))) < 1e-2


def test_r2c_odd_length_mirror():
    x = np.arange(5, dtype=np.float32)
    full = ds.cpu_fft_r2c_full(x)
    ref = np.fft.fft(x.astype(np.float64))
    assert np.allclose(full, ref, atol=1e-5)


def test_magnitude_known_answers():
    s = np.array([3 + 4j, 1 + 0j, 0 + 1j], dtype=np.complex64)
    assert abs(ds.magnitude(s)[0] - 5.0) < 1e-5
    assert abs(ds.magnitude(s, squared=True)[0] - 25.0) < 1e-3
    assert abs(ds.magnitude(s)[1] - 1.0) < 1e-5


def test_empty_input_raises():
    for f in (ds.cpu_fft_c2c, ds.cpu_ifft_c2c):
        with pytest.raises(RuntimeError, match="empty input"):
            f(np.zeros(0, np.complex64))
    with pytest.raises(RuntimeError, match="empty input"):
        ds.cpu_fft_r2c(np.zeros(0, np.float32))


def _kwargs(fn):
    """Keyword names and defaults from a pybind11 signature line."""
    sig = fn.__doc__.splitlines()[0]
    body = sig[sig.index("(") + 1: sig.rindex(")")]
    out = []
    for part in re.split(r",\s*(?![^\[]*\])", body):
        name = part.split(":")[0].strip()
        dflt = part.split("=")[1].strip() if "=" in part else None
        out.append((name, dflt))
    return out[1:] if out and out[0][0] == "self" else out


def test_reference_api_surface():
    for name in ("ROCmGPUContext", "FFTProcessorROCm", "SpectrumMaximaFinderROCm", "ComplexToMagROCm",
                 "cpu_fft_c2c", "cpu_ifft_c2c", "cpu_fft_r2c", "cpu_fft_r2c_full", "magnitude"):
        assert hasattr(ds, name), name
    assert _kwargs(ds.FFTProcessorROCm.process_complex) == [
        ("data", None), ("sample_rate", None), ("beam_count", "0"), ("n_point", "0")]
    assert _kwargs(ds.FFTProcessorROCm.process_mag_phase) == [
        ("data", None), ("sample_rate", None), ("beam_count", "0"), ("n_point", "0"), ("include_freq", "True")]
    assert _kwargs(ds.SpectrumMaximaFinderROCm.process) == [
        ("signal", None), ("n_point", None), ("sample_rate", None), ("antenna_count", "0"), ("repeat_count", "1")]
    assert _kwargs(ds.SpectrumMaximaFinderROCm.find_all_maxima) == [
        ("fft_data", None), ("nFFT", None), ("sample_rate", None), ("beam_count", "0"), ("search_start", "0"),
        ("search_end", "0")]
    assert _kwargs(ds.SpectrumMaximaFinderROCm.find_all_maxima_from_signal) == [
        ("signal", None), ("n_point", None), ("sample_rate", None), ("antenna_count", "0"), ("repeat_count", "1"),
        ("search_start", "0"), ("search_end", "0")]
    assert _kwargs(ds.ComplexToMagROCm.process_magnitude) == [
        ("data", None), ("beam_count", "0"), ("norm_coeff", "1.0")]
    assert isinstance(ds.FFTProcessorROCm.nfft, property)
    assert isinstance(ds.SpectrumMaximaFinderROCm.initialized, property)
    assert callable(ds.FFTProcessorROCm.get_profiling) and callable(ds.SpectrumMaximaFinderROCm.get_params)


def test_maxvalue_dtype_matches_c_abi():
    from spectrum_amd import capi

    assert ds.MAXVALUE_DTYPE == capi.MAXVALUE_DTYPE and ds.MAXVALUE_DTYPE.itemsize == 32
