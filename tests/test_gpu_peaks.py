"""GPU parity of the peak searches through the C ABI.

Fused threshold search (window -> FFT -> |X| -> mean+6std -> local maxima ->
first K ascending -> parabolic): the GPU's peak index lists must equal the
oracle's (rules applied to the FP64 spectrum) except where the deciding
comparison is within tau = 8 * 2^-24 * log2(N) * max|X| of a tie; record
fields must match to FP32 spectral accuracy.
"""
import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def run_fused(ctx, x, nfft, *, K, nsigma=6.0, fs=1.0, window=1, mode=capi.PEAKS_THRESHOLD, s0=0, s1=0,
              refine=capi.REFINE_PARABOLIC):
    B, n_point = x.shape
    d_in = ctx.to_device(x)
    d_pk = ctx.alloc(B * max(K, 1) * 32)
    d_ct = ctx.alloc(B * 8)
    d_g = ctx.alloc(B * 32)
    d_st = ctx.alloc(B * 16)
    ctx.fused_peaks(d_in, B, n_point, nfft, d_pk, d_ct, window=window, fs=fs, peak_mode=mode, nsigma=nsigma,
                    max_peaks=K, s0=s0, s1=s1, refine=refine, gmax=d_g, stats=d_st)
    ctx.sync()
    return (d_pk.download(capi.MAXVALUE_DTYPE, (B, max(K, 1))), d_ct.download(np.uint32, (B, 2)),
            d_g.download(capi.MAXVALUE_DTYPE, (B,)), d_st.download(np.float32, (B, 4)))


def search_range(nfft, mode, s0, s1):
    """The oracle's normalisation of the search range (AllMaxima defaults [1, N/2))."""
    if mode == 2:
        s0 = s0 or 1
        s1 = min(s1 or nfft // 2, nfft - 1)
    elif s0 == 0 and s1 == 0:
        s1 = nfft
    return s0, s1


def check_channels(x, nfft, pk, ct, gm, st, *, K, nsigma=6.0, fs=1.0, window=1, mode=1, s0=0, s1=0, chans=None,
                   chunk=128):
    """Every listed channel against the oracle: index lists and uncapped
    counts under the tie rule, record fields on every bin both lists hold,
    statistics, and the global maximum where it is unique beyond tau."""
    idx = list(range(x.shape[0])) if chans is None else list(chans)
    r0, r1 = search_range(nfft, mode, s0, s1)
    n_exc = 0
    for c0 in range(0, len(idx), chunk):
        sub = idx[c0:c0 + chunk]
        X = orc.spectrum64(x[sub], nfft, window)
        for j, b in enumerate(sub):
            recs, found, g, stats, mags = orc.fused_peaks(X[j], fs, mode=mode, nsigma=nsigma, kcap=K, s0=s0, s1=s1)
            tau = 8 * 2.0 ** -24 * np.log2(nfft) * mags.max()
            thr = stats[2] if mode == 1 else -np.inf
            got = pk[b, : ct[b, 0]]
            ok, why = orc.compare_peak_lists(got["index"], recs["index"], mags, thr, tau, K, found_gpu=ct[b, 1],
                                             found_ref=found, s0=r0, s1=r1)
            assert ok, f"channel {b}: {why}"
            if why != "equal":
                n_exc += 1
            assert ct[b, 0] == min(ct[b, 1], K)
            # record fields on the bins both lists hold
            common, gi, ri = np.intersect1d(got["index"], recs["index"], return_indices=True)
            gg, rr = got[gi], recs[ri]
            ref_x = X[j][common]
            assert np.allclose(gg["magnitude"], np.abs(ref_x), rtol=2e-5, atol=tau), b
            assert np.allclose(gg["real"], ref_x.real, atol=2e-5 * mags.max()), b
            assert np.allclose(gg["imag"], ref_x.imag, atol=2e-5 * mags.max()), b
            big = gg["magnitude"] > 1e-3 * mags.max()
            dph = (gg["phase"] - rr["phase"] + 180) % 360 - 180
            assert np.all(np.abs(dph[big]) < 0.05), b
            assert np.allclose(gg["freq_offset"], rr["freq_offset"], atol=2e-3), b
            assert np.allclose(gg["refined_frequency"], rr["refined_frequency"], rtol=1e-5, atol=2e-3 * fs / nfft), b
            if mode == 1:
                assert abs(st[b, 0] - stats[0]) <= 1e-5 * stats[0] + tau
                assert abs(st[b, 2] - stats[2]) <= 1e-5 * stats[2] + tau
                if np.count_nonzero(np.abs(mags - mags.max()) <= tau) == 1:
                    assert gm[b]["index"] == g["index"]
    return n_exc


@pytest.mark.parametrize("nfft", [256, 1024, 4096, 8192, 16384, 32768, 65536, 131072, 262144])
def test_fused_threshold_sizes(ctx, nfft):
    B = max(4, min(64, (1 << 22) // nfft))
    K = 32
    print(f"[case] fused threshold nfft={nfft} batch={B}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0010 + nfft, B, nfft)
    pk, ct, gm, st = run_fused(ctx, x, nfft, K=K)
    check_channels(x, nfft, pk, ct, gm, st, K=K)


@pytest.mark.parametrize("n_point,nfft", [(3000, 4096), (20000, 32768), (50000, 65536), (100000, 131072)])
def test_fused_zero_padded(ctx, n_point, nfft):
    """Ragged input (n_point < nfft): the Hann window spans the n_point
    samples and the rest of the transform is zero padding, on the short path
    and on the split path (2^17: one channel per column workgroup)."""
    print(f"[case] fused zero padded n_point={n_point} nfft={nfft}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED00C0 + n_point, 4, n_point)
    pk, ct, gm, st = run_fused(ctx, x, nfft, K=32)
    check_channels(x, nfft, pk, ct, gm, st, K=32)


@pytest.mark.parametrize("nfft", [4096, 65536, 131072])
def test_fused_threshold_band(ctx, nfft):
    """Threshold search over bins [s0, s1) only: statistics, local maxima and
    the global maximum of the band (the split path's not-full-band branch)."""
    s0, s1 = 100, nfft // 2 + 37
    print(f"[case] fused threshold band nfft={nfft} s0={s0} s1={s1}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED00D0 + nfft, 6, nfft)
    pk, ct, gm, st = run_fused(ctx, x, nfft, K=32, s0=s0, s1=s1)
    check_channels(x, nfft, pk, ct, gm, st, K=32, s0=s0, s1=s1)
    for b in range(x.shape[0]):
        r = pk[b, : ct[b, 0]]
        assert np.all((r["index"] >= s0) & (r["index"] < s1))
        if ct[b, 1]:
            assert s0 <= gm[b]["index"] < s1


@pytest.mark.parametrize("nfft,dtype", [(4096, capi.DTYPE_C32F16), (65536, capi.DTYPE_C32BF16),
                                        (131072, capi.DTYPE_C32F16)])
def test_fused_half_storage_fp32_compute(ctx, nfft, dtype):
    """16-bit complex storage with a per-channel load scale on the FP32 paths
    (short kernel, split path): the oracle runs on exactly what the kernels
    read, the decoded 16-bit values times the scale, so the same FP32 bar
    applies as for complex64 input."""
    B, K = 4, 32
    print(f"[case] fused half storage nfft={nfft} dtype={dtype}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED00E0 + nfft, B, nfft)
    u = capi.to_half_storage(x * 0.125, dtype)
    scale = np.array([1.0, 0.5, 4.0, 2.0], np.float32)
    d_in, d_sc = ctx.to_device(u), ctx.to_device(scale)
    d_pk, d_ct, d_g, d_st = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32), ctx.alloc(B * 16)
    ctx.fused_peaks(d_in, B, nfft, nfft, d_pk, d_ct, max_peaks=K, dtype=dtype, in_scale=d_sc, gmax=d_g, stats=d_st,
                    window=1, fs=1.0)
    ctx.sync()
    pk, ct = d_pk.download(capi.MAXVALUE_DTYPE, (B, K)), d_ct.download(np.uint32, (B, 2))
    gm, st = d_g.download(capi.MAXVALUE_DTYPE, (B,)), d_st.download(np.float32, (B, 4))
    xr = (capi.from_half_storage(u, dtype) * scale[:, None]).astype(np.complex64)
    check_channels(xr, nfft, pk, ct, gm, st, K=K)
    assert np.all(ct[:, 1] >= 1)


def test_fused_empty_batch_is_a_no_op(ctx):
    d, pk, ct = ctx.alloc(64), ctx.alloc(64), ctx.alloc(64)
    ct.upload(np.full(16, 7, np.uint32))
    for nfft in (4096, 65536):
        ctx.fused_peaks(d, 0, nfft, nfft, pk, ct, max_peaks=2)
    ctx.sync()
    assert np.all(ct.download(np.uint32, (16,)) == 7)


def test_config1_64x4096(ctx):
    """Plain FP32 path at the CPU-runnable shape: 64 x 4096, 3 tones, 10 dB, Hann, K=32."""
    print("[case] config1 64x4096", flush=True)
    x = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0001, 64, 4096)
    pk, ct, gm, st = run_fused(ctx, x, 4096, K=32)
    check_channels(x, 4096, pk, ct, gm, st, K=32)
    assert np.all(ct[:, 1] >= 1)


def test_config2_full_size(ctx):
    """Headline shape 1024 x 65536, K=64: all 1024 channels against the
    oracle, plus global properties (counts within cap, records ascending and
    above threshold)."""
    B, N, K = 1024, 65536, 64
    print("[case] config2 1024x65536", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0002, B, N)
    pk, ct, gm, st = run_fused(ctx, x, N, K=K)
    check_channels(x, N, pk, ct, gm, st, K=K)
    for b in range(B):
        r = pk[b, : ct[b, 0]]
        assert np.all(np.diff(r["index"].astype(np.int64)) > 0)
        assert np.all(r["magnitude"] > st[b, 2] * (1 - 1e-6))
        assert ct[b, 0] <= K and ct[b, 0] >= 1
        assert gm[b]["magnitude"] >= r["magnitude"].max() * (1 - 1e-6)


@pytest.mark.parametrize("nfft", [1024, 65536, 131072])
def test_fused_cap_overflow_keeps_first_k(ctx, nfft):
    """nsigma=0 on noise gives far more peaks than K: the first K ascending are kept."""
    print(f"[case] cap overflow nfft={nfft}", flush=True)
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((4, nfft)) + 1j * rng.standard_normal((4, nfft))).astype(np.complex64)
    pk, ct, gm, st = run_fused(ctx, x, nfft, K=16, nsigma=0.0, window=0)
    assert np.all(ct[:, 1] > 16) and np.all(ct[:, 0] == 16)
    check_channels(x, nfft, pk, ct, gm, st, K=16, nsigma=0.0, window=0)


@pytest.mark.parametrize("nfft", [64, 4096, 65536])
def test_fused_degenerate_inputs(ctx, nfft):
    print(f"[case] degenerate nfft={nfft}", flush=True)
    x = np.zeros((3, nfft), dtype=np.complex64)
    x[1] = 1.0  # constant -> all energy in bin 0
    x[2, 0] = 1.0  # impulse -> exactly flat spectrum
    pk, ct, gm, st = run_fused(ctx, x, nfft, K=8, window=0)
    # zeros: nothing; flat: no strict maxima
    assert ct[0, 1] == 0 and ct[2, 1] == 0
    assert st[0, 0] == 0 and st[2, 1] < 1e-5
    # constant: bin 0 (neighbours N-1 and 1 are circular) clears mean + 6 std = 1 + 6 sqrt(N-1) < N
    assert ct[1, 1] == 1 and pk[1, 0]["index"] == 0 and abs(pk[1, 0]["magnitude"] - nfft) < 1e-3 * nfft
    assert gm[1]["index"] == 0


@pytest.mark.parametrize("nfft", [1024, 8192, 65536, 131072])
def test_all_maxima_from_signal(ctx, nfft):
    """AllMaxima rules on a raw signal (FindAllMaximaFromCPU): default range
    [1, N/2), cap min((s1-s0)/2, 1000), no refinement, uncapped count."""
    print(f"[case] all-maxima from signal nfft={nfft}", flush=True)
    x = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0020 + nfft, 6, nfft)
    s0, s1 = 1, nfft // 2
    cap = min((s1 - s0) // 2, 1000)
    pk, ct, gm, st = run_fused(ctx, x, nfft, K=cap, mode=capi.PEAKS_ALL_MAXIMA, window=0, fs=1000.0)
    X = orc.spectrum64(x, nfft, 0)
    for b in range(6):
        recs, cnt = orc.all_maxima(X[b].astype(np.complex64), 1000.0, s0, s1, cap)
        mags = np.abs(X[b])
        tau = 8 * 2.0 ** -24 * np.log2(nfft) * mags.max()
        ok, why = orc.compare_peak_lists(pk[b, : ct[b, 0]]["index"], recs["index"], mags, -np.inf, tau, cap,
                                         found_gpu=ct[b, 1], found_ref=cnt, s0=s0, s1=s1)
        assert ok, f"beam {b}: {why}"
        got = pk[b, : ct[b, 0]]
        assert np.all(got["freq_offset"] == 0)
        assert np.allclose(got["refined_frequency"], got["index"] * 1000.0 / nfft, rtol=1e-6)


def test_peaks_top_and_all_maxima_on_spectrum(ctx):
    """Reference ONE_PEAK/TWO_PEAKS and AllMaxima kernels on a given spectrum."""
    print("[case] peaks_top / all_maxima on spectrum", flush=True)
    B, N, fs = 16, 2048, 10000.0
    x = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0030, B, N)
    X = orc.spectrum64(x, N, 0).astype(np.complex64)
    d_X = ctx.to_device(X)
    for two in (False, True):
        R = 8 if two else 4
        d_o = ctx.alloc(B * R * 32)
        ctx.peaks_top(d_X, d_o, B, N, N // 4, fs, two_peaks=two)
        ctx.sync()
        got = d_o.download(capi.MAXVALUE_DTYPE, (B, R))
        for b in range(B):
            ref = orc.peaks_top(X[b], N // 4, fs, two)
            assert np.array_equal(got[b]["index"], ref["index"])
            for f in ("magnitude", "real", "imag", "freq_offset", "refined_frequency"):
                assert np.allclose(got[b][f], ref[f], rtol=1e-5, atol=1e-4), f
    s0, s1 = 1, N // 2
    cap = min((s1 - s0) // 2, 1000)
    d_o = ctx.alloc(B * cap * 32)
    d_c = ctx.alloc(B * 4)
    ctx.all_maxima(d_X, d_o, d_c, B, N, fs, s0, s1, cap)
    ctx.sync()
    got = d_o.download(capi.MAXVALUE_DTYPE, (B, cap))
    cnt = d_c.download(np.uint32, (B,))
    for b in range(B):
        ref, rc = orc.all_maxima(X[b], fs, s0, s1, cap)
        assert cnt[b] == rc
        n = min(rc, cap)
        assert np.array_equal(got[b, :n]["index"], ref["index"])
        assert np.allclose(got[b, :n]["magnitude"], ref["magnitude"], rtol=1e-6)


def test_all_maxima_known_answer_on_gpu(ctx):
    n = 1024
    X = np.full((1, n), 0.1, dtype=np.complex64)
    X[0, 50], X[0, 120], X[0, 200] = 100, 80, 60
    d_X = ctx.to_device(X)
    cap = min((n // 2 - 1) // 2, 1000)
    d_o = ctx.alloc(cap * 32)
    d_c = ctx.alloc(4)
    ctx.all_maxima(d_X, d_o, d_c, 1, n, 1000.0, 1, n // 2, cap)
    ctx.sync()
    assert d_c.download(np.uint32, (1,))[0] == 3
    assert list(d_o.download(capi.MAXVALUE_DTYPE, (cap,))[:3]["index"]) == [50, 120, 200]
