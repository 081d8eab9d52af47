"""Reference ONE_PEAK / TWO_PEAKS fused into the transform (spx_fused_peaks_top).

The reference pads, runs hipfftExecC2C and then post_kernel_one_peak /
_two_peaks on the stored spectrum (spectrum_processor_rocm.cpp:179-420,
fft_kernel_sources_rocm.hpp:122-519).  spx_fused_peaks_top takes both
argmaxes inside the transform's launch (nfft <= 2^14), or on the row
transforms of the split path (2^15 .. 2^18: three launches, no spectrum in
HBM, neighbours across row blocks from the blocks' edge rows), or on a stored
chunk spectrum (2^19 and up, and an empty search range).

Each case is checked two ways:
  * against spx_fft_c2c + spx_peaks_top on the same input (the same transform
    arithmetic, so the same records: index exact, values to 1e-6 relative);
  * against the oracle's ONE/TWO_PEAKS rules on the FP64 spectrum: the bin
    may differ only where the FP64 magnitudes are within tau of each other
    (tau = 8 * 2^-24 * log2(nfft) * max|X|, the FP32 transform's error), the
    magnitude within tau, the refined frequency of a dominant peak to 1e-3 bin.
Tones sit on the row-block edges of the split layout (bin k = k2 + N2 * k1
with k2 = 0, N2 - 1, a multiple of 16 and one below) so that the neighbour
lookups across blocks and across the k1 columns are exercised.
"""
import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _signal(B, n_point, nfft, bins, seed, noise=0.01):
    """Channel b: unit tones at the exact bins bins[b] (of nfft) plus complex noise."""
    rng = np.random.default_rng(seed)
    n = np.arange(n_point)
    x = noise * (rng.standard_normal((B, n_point)) + 1j * rng.standard_normal((B, n_point)))
    for b in range(B):
        for j, k in enumerate(bins[b]):
            x[b] += (1.0 - 0.1 * j) * np.exp(2j * np.pi * k * n / nfft)
    return x.astype(np.complex64)


def _edge_bins(nfft):
    """Bins on the split layout's row-block edges (N2 = nfft / 256, 16 rows a block)."""
    n2 = max(nfft // 256, 2)
    return [5 * n2 % nfft, (7 * n2 - 1) % nfft, (3 * n2 + 16) % nfft, (3 * n2 + 15) % nfft, 0, 1000 % nfft]


def _run(ctx, x, nfft, n_point, sr, two, window, fs):
    B = x.shape[0]
    R = 8 if two else 4
    d_in = ctx.to_device(x)
    d_f = ctx.alloc(B * R * 32)
    d_spec = ctx.alloc(B * nfft * 8)
    d_u = ctx.alloc(B * R * 32)
    ctx.fused_peaks_top(d_in, d_f, B, n_point, nfft, sr, fs, two_peaks=two, window=window)
    ctx.fft_c2c(d_in, d_spec, B, n_point, nfft, window)
    ctx.peaks_top(d_spec, d_u, B, nfft, sr, fs, two_peaks=two)
    ctx.sync()
    return d_f.download(capi.MAXVALUE_DTYPE, (B, R)), d_u.download(capi.MAXVALUE_DTYPE, (B, R))


def _check(x, nfft, sr, two, window, fs, got, unf):
    R = 8 if two else 4
    # 1. the unfused GPU chain on the same input
    assert np.array_equal(got["index"], unf["index"]), (got["index"], unf["index"])
    for f in ("real", "imag", "magnitude", "phase", "freq_offset", "refined_frequency"):
        a, b = got[f].astype(np.float64), unf[f].astype(np.float64)
        scale = np.abs(unf["magnitude"]).max() if f in ("real", "imag", "magnitude") else 1.0
        atol = {"phase": 1e-4, "freq_offset": 1e-5, "refined_frequency": 1e-6 * fs}.get(f, 1e-6 * scale)
        assert np.allclose(a, b, rtol=1e-6, atol=atol), f
    # 2. the oracle's rules on the FP64 spectrum
    X = orc.spectrum64(x, nfft, window)
    srr = min(sr, nfft)
    for b in range(x.shape[0]):
        ref = orc.peaks_top(X[b].astype(np.complex64), srr, fs, two)
        mags = np.abs(X[b])
        tau = 8 * 2.0 ** -24 * np.log2(nfft) * mags.max()
        for h in range(R // 4):
            g, r = got[b, 4 * h], ref[4 * h]
            gi, ri = int(g["index"]), int(r["index"])
            if gi != ri:
                assert abs(mags[gi] - mags[ri]) <= tau, (b, h, gi, ri)
            else:
                assert abs(float(g["magnitude"]) - float(r["magnitude"])) <= tau, (b, h)
                if mags[gi] > 0.1 * mags.max():
                    assert abs(float(g["freq_offset"]) - float(r["freq_offset"])) <= 1e-3, (b, h)
                    assert abs(float(g["refined_frequency"]) - float(r["refined_frequency"])) <= 1e-3 * fs / nfft
            # neighbour records: the bins either side of the centre
            assert int(got[b, 4 * h + 2]["index"]) == gi
            if 0 < gi < nfft - 1:
                assert int(got[b, 4 * h + 1]["index"]) == gi - 1
                assert int(got[b, 4 * h + 3]["index"]) == gi + 1


CASES = [
    # nfft, n_point, search_range, two, window, B
    (64, 50, 64, False, 0, 7),
    (1024, 1000, 256, True, 1, 16),
    (4096, 4096, 4096, False, 2, 8),
    (16384, 10000, 1 << 15, True, 0, 8),   # search_range > nfft: clamped
    (16384, 16384, 3, False, 1, 4),        # half = 1: bins 0 and nfft-1
    (1 << 15, 30000, 8192, True, 0, 8),
    (1 << 16, 1 << 16, 1 << 16, False, 1, 8),
    (1 << 16, 1 << 16, 1 << 16, True, 0, 8),
    (1 << 17, 100000, 1000, True, 2, 4),
    (1 << 18, 1 << 18, 1 << 18, True, 0, 4),
    (1 << 16, 1 << 16, 1, False, 0, 3),    # half = 0: both ranges empty
    (1 << 19, 300000, 1 << 19, False, 0, 2),
    (1 << 20, 1 << 20, 4096, True, 1, 2),
]


@pytest.mark.parametrize("nfft,n_point,sr,two,window,B", CASES)
def test_fused_top_vs_unfused_and_oracle(ctx, nfft, n_point, sr, two, window, B):
    print(f"[case] fused top nfft={nfft} n_point={n_point} sr={sr} two={two} win={window} B={B}", flush=True)
    fs = 1.0e6
    eb = _edge_bins(nfft)
    half = min(sr, nfft) // 2
    bins = []
    for b in range(B):
        left = eb[b % len(eb)] % max(half, 1)
        right = nfft - 1 - (eb[(b + 2) % len(eb)] % max(half, 1))
        bins.append([left, right] if (b % 2 == 0 or two) else [right, left])
    x = _signal(B, n_point, nfft, bins, seed=nfft + 7 * sr + B)
    got, unf = _run(ctx, x, nfft, n_point, sr, two, window, fs)
    _check(x, nfft, sr, two, window, fs, got, unf)


def test_fused_top_edge_bins_split(ctx):
    """One dominant tone per channel exactly on each kind of row-block edge,
    both sides, n_point = nfft = 2^16 (N2 = 256): the neighbour records come
    from the edge rows."""
    nfft = 1 << 16
    eb = _edge_bins(nfft)[:4] + [nfft // 2 - 1, 16, 255, 256]
    B = len(eb)
    bins = [[k, nfft - 1 - k] for k in eb]
    print(f"[case] fused top edge bins {eb}", flush=True)
    x = _signal(B, nfft, nfft, bins, seed=5, noise=0.05)
    got, unf = _run(ctx, x, nfft, nfft, nfft, True, 0, 1.0e6)
    _check(x, nfft, nfft, True, 0, 1.0e6, got, unf)
    assert [int(i) for i in got[:, 0]["index"]] == eb
    assert [int(i) for i in got[:, 4]["index"]] == [nfft - 1 - k for k in eb]


def test_fused_top_batch_chunks(ctx):
    """nfft = 2^18 over more channels than fit one launch group of the 2 GiB
    workspace (the split path runs 3 launches per chunk)."""
    nfft, B = 1 << 18, 1100
    print(f"[case] fused top chunks nfft={nfft} B={B}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0400, B, nfft)
    d_in = ctx.to_device(x)
    d_f, d_u = ctx.alloc(B * 4 * 32), ctx.alloc(B * 4 * 32)
    ctx.fused_peaks_top(d_in, d_f, B, nfft, nfft, nfft // 2, 1.0e6)
    sub = 64  # the unfused chain, 64 channels at a time
    d_spec = ctx.alloc(sub * nfft * 8)
    for c0 in range(0, B, sub):
        nb = min(sub, B - c0)
        ctx.fft_c2c(d_in.ptr + c0 * nfft * 8, d_spec, nb, nfft, nfft)
        ctx.peaks_top(d_spec, d_u.ptr + c0 * 4 * 32, nb, nfft, nfft // 2, 1.0e6)
    ctx.sync()
    got = d_f.download(capi.MAXVALUE_DTYPE, (B, 4))
    unf = d_u.download(capi.MAXVALUE_DTYPE, (B, 4))
    assert np.array_equal(got["index"], unf["index"])
    assert np.allclose(got["magnitude"], unf["magnitude"], rtol=1e-6)
    bad = np.argwhere(~np.isclose(got["refined_frequency"], unf["refined_frequency"], rtol=1e-6, atol=1e-3))
    assert len(bad) == 0, [(int(b), int(j), got[b, j], unf[b, j]) for b, j in bad[:4]]


def test_fused_top_rejects(ctx):
    d = ctx.alloc(64)
    with pytest.raises(ValueError):
        ctx.fused_peaks_top(d, d, 1, 100, 1000, 10, 1.0)  # nfft not a power of two
    with pytest.raises(ValueError):
        ctx.fused_peaks_top(d, d, 1, 2000, 1024, 10, 1.0)  # n_point > nfft
    ctx.fused_peaks_top(d, d, 0, 10, 16, 10, 1.0)  # empty batch
