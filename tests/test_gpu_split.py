"""Long channels (N >= 2^15, the split four-step path) against the oracle on
inputs that stress the row-block boundaries.

The split path decides local maxima on the first and last row of each row
block in its finish, from edge rows written by the row kernel, and selects
the first K survivors of an unsorted candidate list with a bitmap of survivor
bins (any K).  Noise with nsigma = 0 puts about N/3 maxima on every channel,
many of them on block boundaries; the local-maxima-only mode has no threshold
at all; degenerate spectra have none or one.  The oracle is the FP64
restatement of the rules (oracle/spx_oracle.c); index lists and counts are
compared under the tie rule of oracle.compare_peak_lists.
"""
import os

import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc
from test_gpu_peaks import check_channels

pytestmark = pytest.mark.gpu

LENGTHS = [1 << 15, 1 << 16, 1 << 17, 1 << 18]


def run(ctx, x, nfft, **kw):
    B = x.shape[0]
    K = kw.pop("K")
    d_in = ctx.to_device(x)
    d_pk, d_ct = ctx.alloc(B * K * 32), ctx.alloc(B * 8)
    d_g, d_st = ctx.alloc(B * 32), ctx.alloc(B * 16)
    ctx.fused_peaks(d_in, B, x.shape[1], nfft, d_pk, d_ct, max_peaks=K, gmax=d_g, stats=d_st, **kw)
    ctx.sync()
    return (d_pk.download(capi.MAXVALUE_DTYPE, (B, K)), d_ct.download(np.uint32, (B, 2)),
            d_g.download(capi.MAXVALUE_DTYPE, (B,)), d_st.download(np.float32, (B, 4)))


@pytest.mark.parametrize("nfft", LENGTHS)
@pytest.mark.parametrize("K", [2048, 5000])
def test_split_noise_many_maxima(ctx, nfft, K):
    """nsigma = 0 on noise: about N/3 maxima per channel, many on row-block
    boundaries; the first K ascending are kept (K = 5000 is past the former
    2048 limit of the long path's selection)."""
    print(f"[case] split noise nfft={nfft} K={K}", flush=True)
    rng = np.random.default_rng(nfft + K)
    x = (rng.standard_normal((3, nfft)) + 1j * rng.standard_normal((3, nfft))).astype(np.complex64)
    pk, ct, gm, st = run(ctx, x, nfft, K=K, nsigma=0.0, window=0)
    assert np.all(ct[:, 1] > K) and np.all(ct[:, 0] == K)
    check_channels(x, nfft, pk, ct, gm, st, K=K, nsigma=0.0, window=0)


@pytest.mark.parametrize("nfft", [1 << 15, 1 << 16, 1 << 18])
def test_split_local_maxima_mode_and_band(ctx, nfft):
    """Local-maxima-only search (no threshold, every maximum is a candidate)
    over a sub-band: the AllMaxima rules of the oracle (non-circular, range
    [s0, s1), count uncapped)."""
    print(f"[case] split all-maxima band nfft={nfft}", flush=True)
    x = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0040 + nfft, 3, nfft)
    s0, s1 = 17, nfft // 2 + 5
    cap = 1000
    pk, ct, gm, st = run(ctx, x, nfft, K=cap, peak_mode=capi.PEAKS_ALL_MAXIMA, window=0, fs=1000.0, s0=s0, s1=s1)
    X = orc.spectrum64(x, nfft, 0)
    for i in range(3):
        recs, cnt = orc.all_maxima(X[i].astype(np.complex64), 1000.0, s0, s1, cap)
        mags = np.abs(X[i])
        tau = 8 * 2.0 ** -24 * np.log2(nfft) * mags.max()
        ok, why = orc.compare_peak_lists(pk[i, : ct[i, 0]]["index"], recs["index"], mags, -np.inf, tau, cap,
                                         found_gpu=ct[i, 1], found_ref=cnt, s0=s0, s1=s1)
        assert ok, f"channel {i}: {why}"
        got = pk[i, : ct[i, 0]]
        assert np.all(got["freq_offset"] == 0)
        assert np.allclose(got["refined_frequency"], got["index"] * 1000.0 / nfft, rtol=1e-6)


@pytest.mark.parametrize("nfft", [1 << 15, 1 << 16])
def test_split_degenerate(ctx, nfft):
    print(f"[case] split degenerate nfft={nfft}", flush=True)
    x = np.zeros((3, nfft), dtype=np.complex64)
    x[1] = 1.0
    x[2, 0] = 1.0
    pk, ct, gm, st = run(ctx, x, nfft, K=8, window=0)
    assert ct[0, 1] == 0 and ct[2, 1] == 0
    assert ct[1, 1] == 1 and pk[1, 0]["index"] == 0 and gm[1]["index"] == 0
    assert abs(pk[1, 0]["magnitude"] - nfft) < 1e-3 * nfft


@pytest.mark.parametrize("nfft,chunk", [(1 << 16, 3), (1 << 15, 5), (1 << 17, 2)])
def test_split_chunks_equal_one_launch_group(ctx, nfft, chunk):
    """A batch split into several launch groups (SPX_FS_CHUNK; the workspace
    bound does this for large batches), including a last group shorter than
    the column kernel's four channels per workgroup, gives the same bytes as
    one group."""
    B = 11
    print(f"[case] split chunks nfft={nfft} chunk={chunk}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0050 + nfft, B, nfft)
    one = run(ctx, x, nfft, K=32)
    old = os.environ.get("SPX_FS_CHUNK")
    os.environ["SPX_FS_CHUNK"] = str(chunk)
    try:
        many = run(ctx, x, nfft, K=32)
    finally:
        if old is None:
            del os.environ["SPX_FS_CHUNK"]
        else:
            os.environ["SPX_FS_CHUNK"] = old
    # records beyond counts[c][0] are not written (the buffers are not cleared)
    (pa, ca, ga, sa), (pb, cb, gb, sb) = one, many
    assert np.array_equal(ca, cb)
    for c in range(B):
        n = ca[c, 0]
        assert pa[c, :n].tobytes() == pb[c, :n].tobytes(), (c, pa[c, :n], pb[c, :n])
    assert ga.tobytes() == gb.tobytes() and sa.tobytes() == sb.tobytes()
    check_channels(x, nfft, pa, ca, ga, sa, K=32)


@pytest.mark.parametrize("nfft", LENGTHS)
def test_split_dense_outputs_vs_fp64(ctx, nfft):
    """Dense spectrum, |X| and {|X|, arg X} of the split path against the FP64
    oracle, with zero padding and a Hann window: relative L2 error at FP32
    level (2e-7 log2 N + 1e-7) and max |dX| / max|X| below 2e-6."""
    B, n_point = 3, nfft - 123
    print(f"[case] split dense nfft={nfft}", flush=True)
    rng = np.random.default_rng(nfft + 7)
    x = (rng.standard_normal((B, n_point)) + 1j * rng.standard_normal((B, n_point))).astype(np.complex64)
    d_in = ctx.to_device(x)
    spec, mag, mp = ctx.alloc(B * nfft * 8), ctx.alloc(B * nfft * 4), ctx.alloc(B * nfft * 8)
    ctx.fft_c2c(d_in, spec, B, n_point, nfft, 1)
    ctx.fft_magnitude(d_in, mag, B, n_point, nfft, 1, squared=False, scale=1.0)
    ctx.fft_mag_phase(d_in, mp, B, n_point, nfft, 1)
    ctx.sync()
    S = spec.download(np.complex64, (B, nfft))
    M = mag.download(np.float32, (B, nfft))
    MP = mp.download(np.float32, (B, nfft, 2))
    ref = orc.spectrum64(x, nfft, 1)
    scale = float(np.abs(ref).max())
    rel = np.linalg.norm(S - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.all(rel <= 2e-7 * np.log2(nfft) + 1e-7), rel
    assert np.max(np.abs(S - ref)) <= 2e-6 * scale
    assert np.max(np.abs(M - np.abs(ref))) <= 2e-6 * scale
    assert np.max(np.abs(MP[..., 0] - np.abs(ref))) <= 2e-6 * scale
    big = np.abs(ref) > 1e-3 * scale
    dphi = np.angle(np.exp(1j * (MP[..., 1] - np.angle(ref))))
    assert np.max(np.abs(dphi[big])) < 1e-4
