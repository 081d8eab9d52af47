"""Lengths 2^19 .. 2^24 (three Stockham passes through HBM, peak search on the
stored spectrum) against the FP64 oracle.

The reference plans any power-of-two nFFT = nextPow2(n_point) * repeat_count
with hipFFT (src/fft_func/src/fft_processor_rocm.cpp:176,665-677); these
lengths are past the split four-step path.  Dense outputs are checked at
FP32 accuracy against the FP64 transform; the threshold and AllMaxima
searches with the tie rule of oracle.compare_peak_lists; the facades
(FFTProcessorROCm / SpectrumMaximaFinderROCm through dsp_spectrum) at
n_point = 200000, repeat_count = 2 (nFFT = 2^19).
"""
import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc
from test_gpu_peaks import check_channels

pytestmark = pytest.mark.gpu


def _tol(nfft):
    return 2e-7 * np.log2(nfft) + 1e-7


@pytest.mark.parametrize("nfft,n_point,win", [(1 << 19, (1 << 19) - 77, 1), (1 << 20, 1 << 20, 0),
                                              (1 << 21, 1500000, 2), (1 << 23, (1 << 23) - 5, 3),
                                              (1 << 24, 1 << 24, 1)])
def test_large_dense_vs_fp64(ctx, nfft, n_point, win):
    """Every radix plan: 2^19 = 8 x 256 x 256, 2^20 = 16 x 256 x 256, 2^21 =
    16 x 512 x 256, 2^23 = 16 x 1024 x 512, 2^24 = 16 x 1024 x 1024 (the
    window evaluated in the first pass from its cosine coefficients)."""
    B = 2 if nfft <= (1 << 21) else 1
    print(f"[case] large dense nfft={nfft} n_point={n_point} win={win}", flush=True)
    rng = np.random.default_rng(nfft + n_point)
    x = (rng.standard_normal((B, n_point)) + 1j * rng.standard_normal((B, n_point))).astype(np.complex64)
    d_in = ctx.to_device(x)
    spec, mag, mp = ctx.alloc(B * nfft * 8), ctx.alloc(B * nfft * 4), ctx.alloc(B * nfft * 8)
    ctx.fft_c2c(d_in, spec, B, n_point, nfft, win)
    ctx.fft_magnitude(d_in, mag, B, n_point, nfft, win, squared=True, scale=0.5)
    ctx.fft_mag_phase(d_in, mp, B, n_point, nfft, win)
    ctx.sync()
    S = spec.download(np.complex64, (B, nfft))
    M = mag.download(np.float32, (B, nfft))
    MP = mp.download(np.float32, (B, nfft, 2))
    ref = orc.spectrum64(x, nfft, win)
    scale = float(np.abs(ref).max())
    rel = np.linalg.norm(S - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.all(rel <= _tol(nfft)), rel
    assert np.max(np.abs(S - ref)) <= 4e-6 * scale
    p = np.abs(ref) ** 2
    assert np.max(np.abs(M - 0.5 * p)) <= 1e-5 * 0.5 * p.max()
    assert np.max(np.abs(MP[..., 0] - np.abs(ref))) <= 4e-6 * scale
    big = np.abs(ref) > 1e-3 * scale
    dphi = np.angle(np.exp(1j * (MP[..., 1] - np.angle(ref))))
    assert np.max(np.abs(dphi[big])) < 1e-4


@pytest.mark.parametrize("nfft,B", [(1 << 19, 4), (1 << 20, 3), (1 << 22, 2), (1 << 24, 1)])
def test_large_threshold_search(ctx, nfft, B):
    """mean + 6 std threshold search, K = 64, Hann, multi-tone input."""
    K = 64
    print(f"[case] large threshold nfft={nfft} B={B}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0300 + nfft, B, nfft)
    d_in = ctx.to_device(x)
    pk, ct, g, st = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32), ctx.alloc(B * 16)
    ctx.fused_peaks(d_in, B, nfft, nfft, pk, ct, max_peaks=K, gmax=g, stats=st, window=1, fs=2.0e6)
    ctx.sync()
    pk = pk.download(capi.MAXVALUE_DTYPE, (B, K))
    ct = ct.download(np.uint32, (B, 2))
    gm = g.download(capi.MAXVALUE_DTYPE, (B,))
    st = st.download(np.float32, (B, 4))
    assert np.all(ct[:, 1] >= 1)
    check_channels(x, nfft, pk, ct, gm, st, K=K, fs=2.0e6, chunk=1)


def test_large_noise_band_and_cap(ctx):
    """Noise with nsigma = 0 over a band: thousands of maxima, the first
    3000 kept in ascending order; and the AllMaxima rules."""
    nfft, B = 1 << 19, 2
    print(f"[case] large noise band nfft={nfft}", flush=True)
    rng = np.random.default_rng(91)
    x = (rng.standard_normal((B, nfft)) + 1j * rng.standard_normal((B, nfft))).astype(np.complex64)
    d_in = ctx.to_device(x)
    K = 3000
    pk, ct, g, st = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32), ctx.alloc(B * 16)
    ctx.fused_peaks(d_in, B, nfft, nfft, pk, ct, max_peaks=K, gmax=g, stats=st, window=0, nsigma=0.0, s0=1000,
                    s1=400000)
    ctx.sync()
    P = pk.download(capi.MAXVALUE_DTYPE, (B, K))
    C = ct.download(np.uint32, (B, 2))
    G = g.download(capi.MAXVALUE_DTYPE, (B,))
    S = st.download(np.float32, (B, 4))
    assert np.all(C[:, 1] > K) and np.all(C[:, 0] == K)
    check_channels(x, nfft, P, C, G, S, K=K, nsigma=0.0, window=0, s0=1000, s1=400000, chunk=1)
    # AllMaxima rules: range [1, N/2), cap 1000, no refinement
    cap = 1000
    pk2, ct2 = ctx.alloc(B * cap * 32), ctx.alloc(B * 8)
    ctx.fused_peaks(d_in, B, nfft, nfft, pk2, ct2, max_peaks=cap, window=0, fs=1000.0,
                    peak_mode=capi.PEAKS_ALL_MAXIMA)
    ctx.sync()
    P2 = pk2.download(capi.MAXVALUE_DTYPE, (B, cap))
    C2 = ct2.download(np.uint32, (B, 2))
    X = orc.spectrum64(x, nfft, 0)
    for b in range(B):
        recs, cnt = orc.all_maxima(X[b].astype(np.complex64), 1000.0, 1, nfft // 2, cap)
        mags = np.abs(X[b])
        tau = 8 * 2.0 ** -24 * np.log2(nfft) * mags.max()
        ok, why = orc.compare_peak_lists(P2[b, : C2[b, 0]]["index"], recs["index"], mags, -np.inf, tau, cap,
                                         found_gpu=C2[b, 1], found_ref=cnt, s0=1, s1=nfft // 2)
        assert ok, f"channel {b}: {why}"
        assert np.all(P2[b, : C2[b, 0]]["freq_offset"] == 0)


@pytest.mark.parametrize("dtype,nfft", [(capi.DTYPE_C32F16, 1 << 19), (capi.DTYPE_C32BF16, 1 << 20)])
def test_large_half_storage(ctx, dtype, nfft):
    """fp16 / bf16 storage with a per-channel load scale: the oracle runs on
    the decoded values times the scale."""
    B, K = 2, 32
    print(f"[case] large half storage dtype={dtype} nfft={nfft}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0310, B, nfft)
    u = capi.to_half_storage(x * 0.125, dtype)
    scale = np.array([2.0, 0.5], np.float32)
    d_in, d_sc = ctx.to_device(u), ctx.to_device(scale)
    pk, ct, g, st = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32), ctx.alloc(B * 16)
    ctx.fused_peaks(d_in, B, nfft, nfft, pk, ct, max_peaks=K, dtype=dtype, in_scale=d_sc, gmax=g,
                    stats=st, window=1, fs=1.0)
    ctx.sync()
    xr = (capi.from_half_storage(u, dtype) * scale[:, None]).astype(np.complex64)
    check_channels(xr, nfft, pk.download(capi.MAXVALUE_DTYPE, (B, K)), ct.download(np.uint32, (B, 2)),
                   g.download(capi.MAXVALUE_DTYPE, (B,)), st.download(np.float32, (B, 4)), K=K, chunk=1)


def test_large_through_facades():
    """nFFT = 2^19 through FFTProcessorROCm (n_point = 300000) and
    SpectrumMaximaFinderROCm (ONE_PEAK, n_point = 200000, repeat_count = 2:
    nextPow2(200000) * 2), as the reference plans them."""
    from spectrum_amd import dsp_spectrum as ds

    print("[case] large facades", flush=True)
    g = ds.ROCmGPUContext(0)
    fs = 1.0e6
    t = np.arange(300000)
    x3 = np.stack([np.exp(2j * np.pi * f0 * t / fs) for f0 in (12345.6, 67890.1)]).astype(np.complex64)
    fft = ds.FFTProcessorROCm(g)
    got = fft.process_complex(x3, sample_rate=fs)
    assert fft.nfft == 1 << 19
    ref = orc.spectrum64(x3, 1 << 19, 0)
    assert np.max(np.abs(got - ref)) <= 4e-6 * np.abs(ref).max()
    n_point = 200000
    x = np.ascontiguousarray(x3[:, :n_point])
    f = ds.SpectrumMaximaFinderROCm(g)
    res = f.process(x, n_point=n_point, sample_rate=fs, repeat_count=2)
    assert f.get_params()["nFFT"] == 1 << 19
    for r, f0 in zip(res, (12345.6, 67890.1)):
        assert abs(r["freq_hz"] - f0) < fs / (1 << 19)


def test_large_launch_groups(ctx):
    """A batch larger than one launch group of the 2 GiB workspace (2^19:
    227 channels per group): channels on both sides of the group boundary
    against the FP64 oracle, for the complex spectrum and the threshold
    search."""
    nfft, B, K = 1 << 19, 300, 32
    print(f"[case] large launch groups nfft={nfft} B={B}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0320, B, nfft)
    d_in = ctx.to_device(x)
    spec = ctx.alloc(B * nfft * 8)
    pk, ct, g, st = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32), ctx.alloc(B * 16)
    ctx.fft_c2c(d_in, spec, B, nfft, nfft, 1)
    ctx.fused_peaks(d_in, B, nfft, nfft, pk, ct, max_peaks=K, gmax=g, stats=st, window=1, fs=1.0)
    ctx.sync()
    idx = np.array([0, 226, 227, 228, 299])
    S = spec.download(np.complex64, (B, nfft))[idx]
    ref = orc.spectrum64(x[idx], nfft, 1)
    rel = np.linalg.norm(S - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.all(rel <= _tol(nfft)), rel
    P = pk.download(capi.MAXVALUE_DTYPE, (B, K))[idx]
    C = ct.download(np.uint32, (B, 2))[idx]
    G = g.download(capi.MAXVALUE_DTYPE, (B,))[idx]
    T = st.download(np.float32, (B, 4))[idx]
    check_channels(x[idx], nfft, P, C, G, T, K=K, chunk=1)
