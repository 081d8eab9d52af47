"""The two long-channel peak-search implementations agree bin for bin.

N >= 2^15 has two GPU implementations of the same search: the three-launch
split path (column kernel, row kernel, per-channel finish) and the single
persistent launch.  SPX_LARGE_PATH=split|fused selects one; both are checked
against each other on inputs that stress the split path's block-boundary
rows (noise with nsigma = 0 puts thousands of local maxima on them; the
local-maxima-only mode has no threshold at all), on band-limited searches
and on degenerate spectra.  The oracle checks of each path's default
lengths live in test_gpu_peaks.py.
"""
import os

import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def run(ctx, x, nfft, path, **kw):
    B = x.shape[0]
    K = kw.pop("K")
    d_in = ctx.to_device(x)
    d_pk, d_ct = ctx.alloc(B * K * 32), ctx.alloc(B * 8)
    d_g, d_st = ctx.alloc(B * 32), ctx.alloc(B * 16)
    old = os.environ.get("SPX_LARGE_PATH")
    os.environ["SPX_LARGE_PATH"] = path
    try:
        ctx.fused_peaks(d_in, B, x.shape[1], nfft, d_pk, d_ct, max_peaks=K, gmax=d_g, stats=d_st, **kw)
        ctx.sync()
    finally:
        if old is None:
            del os.environ["SPX_LARGE_PATH"]
        else:
            os.environ["SPX_LARGE_PATH"] = old
    return (d_pk.download(capi.MAXVALUE_DTYPE, (B, K)), d_ct.download(np.uint32, (B, 2)),
            d_g.download(capi.MAXVALUE_DTYPE, (B,)), d_st.download(np.float32, (B, 4)))


def assert_same(a, b, *, stats=True):
    pa, ca, ga, sa = a
    pb, cb, gb, sb = b
    assert np.array_equal(ca, cb), (ca[:4], cb[:4])
    for i in range(ca.shape[0]):
        n = ca[i, 0]
        assert np.array_equal(pa[i, :n]["index"], pb[i, :n]["index"]), i
        for f in ("real", "imag", "magnitude", "freq_offset"):
            assert np.allclose(pa[i, :n][f], pb[i, :n][f], rtol=1e-5, atol=1e-6 * max(1.0, float(sa[i, 0]))), (i, f)
    assert np.array_equal(ga["index"], gb["index"])
    if stats:
        assert np.allclose(sa, sb, rtol=1e-5, atol=1e-7)


LENGTHS = [1 << 15, 1 << 16, 1 << 17, 1 << 18]


@pytest.mark.parametrize("nfft", LENGTHS)
def test_split_equals_fused_tones(ctx, nfft):
    B = max(4, (1 << 22) // nfft)
    print(f"[case] split/fused tones nfft={nfft} batch={B}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0030 + nfft, B, nfft)
    assert_same(run(ctx, x, nfft, "split", K=64), run(ctx, x, nfft, "fused", K=64))


@pytest.mark.parametrize("nfft", LENGTHS)
def test_split_equals_fused_noise_many_maxima(ctx, nfft):
    """nsigma = 0 on noise: about N/3 maxima per channel, many of them on
    row-block boundaries; K = 2048 keeps the first 2048 ascending."""
    print(f"[case] split/fused noise nfft={nfft}", flush=True)
    rng = np.random.default_rng(nfft)
    x = (rng.standard_normal((3, nfft)) + 1j * rng.standard_normal((3, nfft))).astype(np.complex64)
    a = run(ctx, x, nfft, "split", K=2048, nsigma=0.0, window=0)
    b = run(ctx, x, nfft, "fused", K=2048, nsigma=0.0, window=0)
    assert np.all(a[1][:, 1] > 2048)
    assert_same(a, b)


@pytest.mark.parametrize("nfft", [1 << 15, 1 << 16])
def test_split_local_maxima_mode_and_band(ctx, nfft):
    """Local-maxima-only search (no threshold, every maximum is a candidate)
    over a sub-band, checked against the oracle's rules on the FP64 spectrum."""
    print(f"[case] split all-maxima band nfft={nfft}", flush=True)
    x = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0040 + nfft, 3, nfft)
    s0, s1 = 17, nfft // 2 + 5
    cap = 1000
    a = run(ctx, x, nfft, "split", K=cap, peak_mode=capi.PEAKS_ALL_MAXIMA, window=0, fs=1000.0, s0=s0, s1=s1)
    b = run(ctx, x, nfft, "fused", K=cap, peak_mode=capi.PEAKS_ALL_MAXIMA, window=0, fs=1000.0, s0=s0, s1=s1)
    assert_same(a, b)
    X = orc.spectrum64(x, nfft, 0)
    for i in range(3):
        recs, cnt = orc.all_maxima(X[i].astype(np.complex64), 1000.0, s0, s1, cap)
        n = a[1][i, 0]
        assert n == min(cnt, cap) and a[1][i, 1] == cnt
        assert np.array_equal(a[0][i, :n]["index"], recs["index"][:n])


@pytest.mark.parametrize("nfft", [1 << 15, 1 << 16])
def test_split_degenerate(ctx, nfft):
    print(f"[case] split degenerate nfft={nfft}", flush=True)
    x = np.zeros((3, nfft), dtype=np.complex64)
    x[1] = 1.0
    x[2, 0] = 1.0
    pk, ct, gm, st = run(ctx, x, nfft, "split", K=8, window=0)
    assert ct[0, 1] == 0 and ct[2, 1] == 0
    assert ct[1, 1] == 1 and pk[1, 0]["index"] == 0 and gm[1]["index"] == 0
    assert_same((pk, ct, gm, st), run(ctx, x, nfft, "fused", K=8, window=0))


@pytest.mark.parametrize("nfft,chunk", [(1 << 16, 3), (1 << 15, 5), (1 << 17, 2)])
def test_split_chunks_equal_one_launch_group(ctx, nfft, chunk):
    """A batch split into several launch groups (SPX_FS_CHUNK; the workspace
    bound does this for large batches), including a last group shorter than
    the column kernel's four channels per workgroup, gives the same bytes as
    one group."""
    B = 11
    print(f"[case] split chunks nfft={nfft} chunk={chunk}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0050 + nfft, B, nfft)
    one = run(ctx, x, nfft, "split", K=32)
    old = os.environ.get("SPX_FS_CHUNK")
    os.environ["SPX_FS_CHUNK"] = str(chunk)
    try:
        many = run(ctx, x, nfft, "split", K=32)
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


@pytest.mark.parametrize("nfft", LENGTHS)
def test_split_dense_outputs_equal_single_launch(ctx, nfft):
    """Dense spectrum, |X| and {|X|, arg X} of the split path against the
    single persistent launch, with zero padding and a window."""
    B, n_point = 3, nfft - 123
    print(f"[case] split dense nfft={nfft}", flush=True)
    rng = np.random.default_rng(nfft + 7)
    x = (rng.standard_normal((B, n_point)) + 1j * rng.standard_normal((B, n_point))).astype(np.complex64)
    d_in = ctx.to_device(x)
    outs = {}
    for path in ("split", "fused"):
        old = os.environ.get("SPX_LARGE_PATH")
        os.environ["SPX_LARGE_PATH"] = path
        try:
            spec, mag, mp = ctx.alloc(B * nfft * 8), ctx.alloc(B * nfft * 4), ctx.alloc(B * nfft * 8)
            ctx.fft_c2c(d_in, spec, B, n_point, nfft, 1)
            ctx.fft_magnitude(d_in, mag, B, n_point, nfft, 1, squared=False, scale=1.0)
            ctx.fft_mag_phase(d_in, mp, B, n_point, nfft, 1)
            ctx.sync()
            outs[path] = (spec.download(np.complex64, (B, nfft)), mag.download(np.float32, (B, nfft)),
                          mp.download(np.float32, (B, nfft, 2)))
        finally:
            if old is None:
                del os.environ["SPX_LARGE_PATH"]
            else:
                os.environ["SPX_LARGE_PATH"] = old
    (sa, ma, pa), (sb, mb, pb) = outs["split"], outs["fused"]
    scale = float(np.abs(sb).max())
    assert np.max(np.abs(sa - sb)) <= 2e-6 * scale
    assert np.max(np.abs(ma - mb)) <= 2e-6 * scale
    assert np.max(np.abs(pa[..., 0] - pb[..., 0])) <= 2e-6 * scale
    big = pb[..., 0] > 1e-3 * scale
    dphi = np.angle(np.exp(1j * (pa[..., 1] - pb[..., 1])))
    assert np.max(np.abs(dphi[big])) < 1e-4
