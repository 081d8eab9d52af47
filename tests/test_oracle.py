"""CPU tests that pin the oracle before anything is compared against it.

Every oracle function is checked against something that does not share its
code: a brute-force long-double DFT, numpy's FFT, closed forms, the
reference's own CPU FFT compiled from its sources (oracle/_ref), and the
reference tests' known answers.  Planted-mistake tests show each check fails
for a plausible bug (conjugated twiddles, a dropped term, a shifted index).
"""
import numpy as np
import pytest

from oracle import oracle as orc

RNG = np.random.default_rng(1234)


def crandn(*shape):
    return (RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)).astype(np.complex128)


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [2 ** k for k in range(1, 13)])
def test_fft_vs_bruteforce(n):
    x = crandn(n)
    got = orc.fft64(x)
    ref = orc.dft_brute(x)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.sqrt(n) * np.max(np.abs(ref))


@pytest.mark.parametrize("n", [2 ** k for k in range(1, 19)])
def test_fft_vs_numpy(n):
    x = crandn(n)
    got = orc.fft64(x)
    ref = np.fft.fft(x)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.log2(max(n, 2)) * np.linalg.norm(ref)


def integer_tone_checks(fft, n=1024, k0=37, amp=1.5, ph=0.7):
    """X[k0] = A N e^{j phi}, zero elsewhere (closed form)."""
    t = np.arange(n)
    x = amp * np.exp(1j * (2 * np.pi * k0 * t / n + ph))
    X = fft(x)
    ok_peak = abs(X[k0] - amp * n * np.exp(1j * ph)) < 1e-9 * n
    rest = np.delete(X, k0)
    ok_rest = np.max(np.abs(rest)) < 1e-9 * n
    return ok_peak and ok_rest


def impulse_checks(fft, n=256, m=3):
    """FFT of delta[n - m] is W_N^{m k} (the twiddles themselves)."""
    x = np.zeros(n, dtype=np.complex128)
    x[m] = 1.0
    k = np.arange(n)
    return np.max(np.abs(fft(x) - np.exp(-2j * np.pi * m * k / n))) < 1e-12


def parseval_checks(fft, n=4096):
    x = crandn(n)
    X = fft(x)
    return abs(np.sum(np.abs(X) ** 2) - n * np.sum(np.abs(x) ** 2)) < 1e-9 * n * np.sum(np.abs(x) ** 2)


def test_fft_closed_forms():
    assert integer_tone_checks(orc.fft64)
    assert impulse_checks(orc.fft64)
    assert parseval_checks(orc.fft64)
    # the transform is unnormalised: a constant gives N at bin 0
    X = orc.fft64(np.ones(64))
    assert abs(X[0] - 64) < 1e-12 and np.max(np.abs(X[1:])) < 1e-12


def test_closed_forms_catch_planted_mistakes():
    # conjugated twiddles (inverse instead of forward) -> tone lands at -k0
    assert not integer_tone_checks(lambda x: np.fft.ifft(x) * len(x))
    assert not impulse_checks(lambda x: np.fft.ifft(x) * len(x))
    # a dropped term (first sample ignored)
    def dropped(x):
        y = np.array(x, dtype=np.complex128)
        y[0] = 0
        return np.fft.fft(y)
    assert not impulse_checks(lambda x: dropped(np.roll(x, -3)) if False else dropped(x), m=0)
    # off-by-one bin (output rotated)
    assert not integer_tone_checks(lambda x: np.roll(np.fft.fft(x), 1))
    # normalised transform (1/N)
    assert not parseval_checks(lambda x: np.fft.fft(x) / len(x))


@pytest.mark.skipif(not orc.have_reference(), reason="oracle/_ref (reference CpuFFT) not built")
@pytest.mark.parametrize("n", [16, 512, 1024, 4096, 65536])
def test_oracle_vs_reference_cpufft(n):
    """The reference's CpuFFT::ForwardC2C (pocketfft float) agrees with the FP64
    oracle to the reference's own tolerance max|dX|/nFFT < 1e-5
    (tests/test_fft_cpu_reference_rocm.hpp:155-157)."""
    x = crandn(4, n).astype(np.complex64)
    ref = orc.ref_cpufft(x)
    ora = orc.spectrum64(x, n, 0)
    assert np.max(np.abs(ref - ora)) / n < 1e-5
    # and much tighter in the relative L2 sense
    assert np.linalg.norm(ref - ora) / np.linalg.norm(ora) < 1e-6
    mt = orc.ref_pocketfft_mt(x, 2)
    assert np.array_equal(mt, ref)


@pytest.mark.skipif(not orc.have_reference(), reason="oracle/_ref (reference CpuFFT) not built")
def test_reference_cpufft_multibeam_tones():
    """4 beams, N=512, tones at 50/125/200/375 Hz, fs 1000: reference test
    tests/test_fft_cpu_reference_rocm.hpp:231-286 (same tolerance)."""
    n, fs = 512, 1000.0
    t = np.arange(n) / fs
    x = np.stack([np.exp(2j * np.pi * f * t) for f in (50, 125, 200, 375)]).astype(np.complex64)
    ref = orc.ref_cpufft(x)
    ora = orc.spectrum64(x, n, 0)
    assert np.max(np.abs(ref - ora)) / n < 1e-5


# ------------------------------------------------------------------ windows
def test_windows_match_reference_formula():
    n = 1000
    p = np.arange(n)
    th = 2 * np.pi * p / (n - 1)
    assert np.allclose(orc.window(1, n), 0.5 * (1 - np.cos(th)), atol=1e-15)
    assert np.allclose(orc.window(2, n), 0.54 - 0.46 * np.cos(th), atol=1e-15)
    assert np.allclose(orc.window(3, n), 0.42 - 0.5 * np.cos(th) + 0.08 * np.cos(2 * th), atol=1e-15)
    assert np.allclose(orc.window(0, n), 1.0)
    # symmetric Hann equals numpy's symmetric Hann
    assert np.allclose(orc.window(1, n), np.hanning(n), atol=1e-15)
    assert np.all(orc.window(1, 1) == 1.0)


# ------------------------------------------------------------------ known answers
def test_known_answer_peak_bin_102():
    """100 Hz at fs 1000, N 1024 -> peak bin 102 (tests/test_fft_processor_rocm.hpp:59-83)."""
    n, fs = 1024, 1000.0
    x = np.exp(2j * np.pi * 100.0 * np.arange(n) / fs).astype(np.complex64)
    X = orc.spectrum64(x, n, 0)[0]
    assert int(np.argmax(np.abs(X[: n // 2]))) == 102
    _, _, g, _, _ = orc.fused_peaks(X, fs)
    assert g["index"] == 102


def test_known_answer_all_maxima_50_120_200():
    """0.1 floor with spikes at 50/120/200 -> exactly {50, 120, 200}
    (tests/test_spectrum_maxima_rocm.hpp:179-206), range [1, N/2)."""
    n = 1024
    X = np.full(n, 0.1, dtype=np.complex64)
    X[50], X[120], X[200] = 100, 80, 60
    recs, cnt = orc.all_maxima(X, 1000.0, 1, n // 2, min((n // 2 - 1) // 2, 1000))
    assert cnt == 3 and list(recs["index"]) == [50, 120, 200]
    assert np.allclose(recs["refined_frequency"], np.array([50, 120, 200]) * 1000.0 / n)
    assert np.all(recs["freq_offset"] == 0)
    r2, found, _, _, _ = orc.fused_peaks(X, 1000.0, mode=2, kcap=1000)
    assert list(r2["index"]) == [50, 120, 200] and found == 3


def test_known_answer_magnitudes():
    X = np.array([3 + 4j, 1 + 0j, 0 + 1j], dtype=np.complex64)
    assert np.allclose(np.abs(X), [5, 1, 1])  # python/t_cpu_fft.py:112-121
    r = orc.peaks_top(np.array([0, 3 + 4j, 0, 0], dtype=np.complex64), 4, 4.0)
    assert r[0]["magnitude"] == 5.0 and r[0]["index"] == 1


def test_frequency_axis():
    """freq[k] = k * fs / N (tests/test_fft_processor_rocm.hpp:174-207)."""
    n, fs = 256, 48000.0
    X = np.zeros(n, dtype=np.complex64)
    X[10] = 5
    X[9] = X[11] = 1
    r, _ = orc.all_maxima(X, fs, 1, n // 2, 10)
    assert abs(r[0]["refined_frequency"] - 10 * fs / n) < 1e-3


# ------------------------------------------------------------------ peak rules
def quad_spectrum(n, k0, x0, a=-3.0, c=50.0):
    X = np.full(n, 1.0, dtype=np.complex64)
    for d in (-1, 0, 1):
        X[k0 + d] = a * (d - x0) ** 2 + c
    return X


@pytest.mark.parametrize("x0", [-0.4, -0.1, 0.0, 0.25, 0.45])
def test_parabolic_exact_on_quadratic(x0):
    """Parabolic refinement is exact on quadratic samples."""
    n, k0 = 64, 20
    X = quad_spectrum(n, k0, x0)
    recs, found, g, _, _ = orc.fused_peaks(X, 64.0, nsigma=1.0, kcap=4)
    assert found == 1 and recs[0]["index"] == k0
    assert abs(recs[0]["freq_offset"] - x0) < 1e-6
    assert abs(recs[0]["refined_frequency"] - (k0 + x0)) < 1e-4
    top = orc.peaks_top(X, n, 64.0)  # search_range n: left half [0, n/2)
    assert top[0]["index"] == k0 and abs(top[0]["freq_offset"] - x0) < 1e-5


def test_parabolic_catches_sign_mistake():
    X = quad_spectrum(64, 20, 0.3)
    recs, *_ = orc.fused_peaks(X, 64.0, nsigma=1.0, kcap=4)
    wrong = -recs[0]["freq_offset"]  # a flipped (yl - yr) sign
    assert abs(wrong - 0.3) > 0.5


def test_fused_rules_constant_and_single_tone():
    n = 4096
    recs, found, _, st, _ = orc.fused_peaks(np.ones(n, dtype=np.complex64), 1.0)
    assert found == 0 and st[1] == 0.0
    # noise-free integer-bin tone through Hann: exactly one peak above mean + 6 std
    x = np.exp(2j * np.pi * 300 * np.arange(n) / n).astype(np.complex64)
    X = orc.spectrum64(x, n, 1)[0]
    recs, found, g, st, _ = orc.fused_peaks(X, 1.0)
    assert found == 1 and recs[0]["index"] == 300 and g["index"] == 300
    # std is the population std of |X|
    m = np.abs(X)
    assert abs(st[0] - m.mean()) < 1e-9 * m.max() and abs(st[1] - m.std()) < 1e-6 * m.max()


def test_fused_rules_circular_neighbours_and_cap():
    n = 32
    m = np.ones(n)
    m[0] = 10  # peak at bin 0 needs the circular neighbour n-1
    m[5::4] = 5
    X = m.astype(np.complex64)
    recs, found, *_ = orc.fused_peaks(X, 1.0, nsigma=0.0, kcap=3)
    assert found == 1 + len(range(5, n, 4)) and list(recs["index"]) == [0, 5, 9]


def test_threshold_is_strict():
    n = 16
    X = np.zeros(n, dtype=np.complex64)
    X[4] = 2.0
    recs, found, _, st, _ = orc.fused_peaks(X, 1.0, nsigma=0.0)
    assert found == 1
    # threshold equal to the peak value rejects it (strict >)
    recs, found, _, st, _ = orc.fused_peaks(X, 1.0, nsigma=(2.0 - st[0]) / st[1])
    assert found == 0


# ------------------------------------------------------------------ reference ONE/TWO_PEAKS rules
def test_one_peak_rules():
    n, fs = 64, 640.0
    X = np.zeros(n, dtype=np.complex64)
    X[5], X[4], X[6] = 10, 4, 6        # left range [0, 8)
    X[60], X[59], X[61] = 12, 5, 5     # right range [56, 64)
    r = orc.peaks_top(X, 16, fs)
    # right side is larger -> center 60, mirrored frequency fs - (k + d) fs / N
    assert r[0]["index"] == 60 and abs(r[0]["refined_frequency"] - (fs - 60 * fs / n)) < 1e-3
    assert r[1]["index"] == 59 and r[3]["index"] == 61
    r2 = orc.peaks_top(X, 16, fs, two_peaks=True)
    assert r2[0]["index"] == 5 and r2[4]["index"] == 60
    d = 0.5 * (4 - 6) / (4 - 20 + 6)
    assert abs(r2[0]["freq_offset"] - d) < 1e-6
    assert abs(r2[0]["refined_frequency"] - (5 + d) * fs / n) < 1e-3
    # left wins ties
    X[60] = 10
    assert orc.peaks_top(X, 16, fs)[0]["index"] == 5


def test_all_maxima_cap_keeps_lowest():
    n = 256
    X = np.zeros(n, dtype=np.complex64)
    X[10:120:10] = np.arange(1, 12)  # 11 maxima, increasing
    recs, cnt = orc.all_maxima(X, 1.0, 1, n // 2, 4)
    assert cnt == 11 and list(recs["index"]) == [10, 20, 30, 40]


# ------------------------------------------------------------------ comparator
def test_tie_tolerant_comparator():
    m = np.array([0, 5, 0, 3, 3.0000001, 0, 2, 0], dtype=np.float64)
    ok, _ = orc.compare_peak_lists([1, 4, 6], [1, 4, 6], m, 1.0, 1e-3, 10)
    assert ok
    ok, why = orc.compare_peak_lists([1, 6], [1, 4, 6], m, 1.0, 1e-3, 10)
    assert ok, why  # bin 4 vs 3 is a near tie
    ok, _ = orc.compare_peak_lists([6], [1, 6], m, 1.0, 1e-3, 10)
    assert not ok  # bin 1 is a clear peak


def _capped_spectrum():
    # maxima at 10, 20, 30, 33, 40, 43, 50; bin 20 is a near tie with bin 21
    m = np.zeros(64)
    for k, v in ((10, 5.0), (20, 4.0), (30, 6.0), (33, 6.0), (40, 7.0), (43, 7.0), (50, 8.0)):
        m[k] = v
    m[21] = 4.0 - 1e-6
    return m


def test_comparator_cap_shift_needs_matching_prefix():
    """At the cap an excused flip may shift the tail, but every non-excused
    bin up to the smaller last bin must agree: the lists below differ at 30,
    33, 40 and 43, which are clear decisions, and must fail."""
    m = _capped_spectrum()
    ok, why = orc.compare_peak_lists([10, 20, 33, 43], [10, 30, 40, 50], m, 1.0, 1e-3, 4)
    assert not ok, why
    # the legitimate form of the same flip: gpu keeps 20 and loses 50 past the cap
    ok, why = orc.compare_peak_lists([10, 20, 30, 33], [10, 30, 33, 40], m, 1.0, 1e-3, 4)
    assert ok, why
    # a full list against a complete shorter one: the shorter one holds every
    # peak, so a clear bin below the full list's last bin must be in both
    ok, _ = orc.compare_peak_lists([10, 30, 33, 40], [10, 30, 40], m, 1.0, 1e-3, 4)
    assert not ok  # 33 is missing from the complete reference list
    ok, why = orc.compare_peak_lists([10, 20, 30, 33], [10, 30, 33], m, 1.0, 1e-3, 4)
    assert ok, why  # the only difference is the excused 20


def test_comparator_found_counts_under_tie_rule():
    m = _capped_spectrum()
    # bins 20 and 21 are within tau of each other, so either may be the
    # maximum there: the counts may differ by up to two, not by three
    ok, why = orc.compare_peak_lists([10, 30], [10, 30], m, 1.0, 1e-3, 2, found_gpu=8, found_ref=6)
    assert ok, why
    ok, _ = orc.compare_peak_lists([10, 30], [10, 30], m, 1.0, 1e-3, 2, found_gpu=9, found_ref=6)
    assert not ok
    ok, _ = orc.compare_peak_lists([10, 30], [10, 30], m, 1.0, 1e-3, 2, found_gpu=6, found_ref=6)
    assert ok
