"""Host-resident entry points (spx_*_host): the batch streamed in chunks
through pinned staging on three streams gives the device entry points'
results bit for bit, for pageable and pinned caller buffers, with several
chunks and a short last chunk."""
import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def device_dense(ctx, x, nfft, window, kind):
    B, n = x.shape
    d_in = ctx.to_device(x)
    per = 4 if kind == "mag" else 8
    out = ctx.alloc(B * nfft * per)
    if kind == "c2c":
        ctx.fft_c2c(d_in, out, B, n, nfft, window)
        ctx.sync()
        return out.download(np.complex64, (B, nfft))
    if kind == "magphase":
        ctx.fft_mag_phase(d_in, out, B, n, nfft, window)
        ctx.sync()
        return out.download(np.float32, (B, nfft, 2))
    ctx.fft_magnitude(d_in, out, B, n, nfft, window, squared=False, scale=0.5)
    ctx.sync()
    return out.download(np.float32, (B, nfft))


@pytest.mark.parametrize("B,n,nfft,chunk_bytes", [(37, 1000, 1024, 5 * 8192), (9, 65536, 65536, 2 * 524288),
                                                 (5, 3000, 4096, 1 << 30), (3, 3000, 4096, 1 << 62)])
def test_host_dense_equals_device(ctx, B, n, nfft, chunk_bytes):
    print(f"[case] host dense B={B} n={n} nfft={nfft} chunk={chunk_bytes}", flush=True)
    rng = np.random.default_rng(B * 7 + n)
    x = (rng.standard_normal((B, n)) + 1j * rng.standard_normal((B, n))).astype(np.complex64)
    ctx.set_host_chunk(chunk_bytes)
    try:
        got = ctx.fft_c2c_host(x, nfft, window=1)
        assert got.tobytes() == device_dense(ctx, x, nfft, 1, "c2c").tobytes()
        got = ctx.fft_mag_phase_host(x, nfft, window=0)
        assert got.tobytes() == device_dense(ctx, x, nfft, 0, "magphase").tobytes()
        got = ctx.fft_magnitude_host(x, nfft, window=2, squared=False, scale=0.5)
        assert got.tobytes() == device_dense(ctx, x, nfft, 2, "mag").tobytes()
    finally:
        ctx.set_host_chunk(32 << 20)


def test_host_c2c_against_oracle(ctx):
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0060, 6, 4096)
    ctx.set_host_chunk(2 * 4096 * 8)
    try:
        got = ctx.fft_c2c_host(x, 4096)
    finally:
        ctx.set_host_chunk(32 << 20)
    ref = orc.spectrum64(x, 4096, 0)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err


def device_peaks(ctx, x, nfft, K, **kw):
    B = x.shape[0]
    d_in = ctx.to_device(x)
    pk, ct, g, st = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32), ctx.alloc(B * 16)
    ctx.fused_peaks(d_in, B, x.shape[1], nfft, pk, ct, max_peaks=K, gmax=g, stats=st, **kw)
    ctx.sync()
    return (pk.download(capi.MAXVALUE_DTYPE, (B, K)), ct.download(np.uint32, (B, 2)),
            g.download(capi.MAXVALUE_DTYPE, (B,)), st.download(np.float32, (B, 4)))


def assert_same_records(a, b):
    pa, ca, ga, sa = a
    pb, cb, gb, sb = b
    assert np.array_equal(ca, cb)
    for i in range(ca.shape[0]):
        n = ca[i, 0]
        assert pa[i, :n].tobytes() == pb[i, :n].tobytes(), i
    assert ga.tobytes() == gb.tobytes() and sa.tobytes() == sb.tobytes()


@pytest.mark.parametrize("B,nfft,K,chunk", [(21, 4096, 32, 4), (13, 65536, 64, 3), (3, 16384, 16, 64)])
def test_host_fused_peaks_equals_device(ctx, B, nfft, K, chunk):
    print(f"[case] host peaks B={B} nfft={nfft} chunk={chunk}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0070 + nfft, B, nfft)
    ctx.set_host_chunk(chunk * nfft * 8)
    try:
        got = ctx.fused_peaks_host(x, nfft, max_peaks=K, window=1, fs=1000.0)
    finally:
        ctx.set_host_chunk(32 << 20)
    assert_same_records(got, device_peaks(ctx, x, nfft, K, window=1, fs=1000.0))
    assert np.all(got[1][:, 1] >= 1)


def test_host_all_maxima_mode_band(ctx):
    nfft, B = 8192, 7
    x = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0080, B, nfft)
    ctx.set_host_chunk(3 * nfft * 8)
    try:
        got = ctx.fused_peaks_host(x, nfft, max_peaks=100, window=0, fs=1000.0,
                                   peak_mode=capi.PEAKS_ALL_MAXIMA, s0=10, s1=3000)
    finally:
        ctx.set_host_chunk(32 << 20)
    ref = device_peaks(ctx, x, nfft, 100, window=0, fs=1000.0, peak_mode=capi.PEAKS_ALL_MAXIMA, s0=10, s1=3000)
    assert_same_records(got, ref)


def test_host_pinned_buffers(ctx):
    """Caller buffers in pinned memory (spx_host_alloc) are transferred by DMA
    without staging; pageable input with pinned output mixes both ways."""
    B, nfft = 11, 8192
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0090, B, nfft)
    xp = ctx.pinned((B, nfft), np.complex64)
    xp.array[...] = x
    outp = ctx.pinned((B, nfft), np.complex64)
    ref = device_dense(ctx, x, nfft, 1, "c2c").tobytes()
    ctx.set_host_chunk(4 * nfft * 8)
    try:
        got = ctx.fft_c2c_host(xp.array, nfft, window=1, out=outp.array)
        assert got.ctypes.data == outp.ptr and outp.array.tobytes() == ref
        outp.array[...] = 0
        ctx.fft_c2c_host(x, nfft, window=1, out=outp.array)
        assert outp.array.tobytes() == ref
        assert ctx.fft_c2c_host(xp.array, nfft, window=1).tobytes() == ref
    finally:
        ctx.set_host_chunk(32 << 20)
        xp.free()
        outp.free()


def test_host_errors_leave_context_usable(ctx):
    x = np.ones((2, 100), np.complex64)
    with pytest.raises(ValueError):
        ctx.fft_c2c_host(x, 96)  # nfft < n_point
    with pytest.raises(ValueError):
        ctx.fft_c2c_host(x, 100)  # not a power of two
    got = ctx.fft_c2c_host(x, 128)
    assert got.shape == (2, 128) and abs(got[0, 0] - 100) < 1e-4


def test_failed_allocation_leaves_context_usable(ctx):
    """A device allocation larger than the GPU fails with SPX_ENOMEM (status
    2); the runtime error it leaves is cleared, so the next launches succeed
    with the same results as before."""
    B, nfft, K = 4, 4096, 16
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED00B0, B, nfft)
    before = device_peaks(ctx, x, nfft, K, window=1, fs=1000.0)
    with pytest.raises(capi.SpxError, match="status 2"):
        ctx.alloc(1 << 52)
    assert_same_records(device_peaks(ctx, x, nfft, K, window=1, fs=1000.0), before)
    assert_same_records(ctx.fused_peaks_host(x, nfft, max_peaks=K, window=1, fs=1000.0), before)


@pytest.mark.parametrize("nfft,dtype", [(4096, capi.DTYPE_C32F16), (65536, capi.DTYPE_C32BF16)])
def test_host_half_storage_with_scale(ctx, nfft, dtype):
    """16-bit complex input (4 bytes per sample) and a per-channel load
    scale, both sliced per chunk: same records as the device entry point."""
    B, K = 9, 32
    print(f"[case] host half storage nfft={nfft} dtype={dtype}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED00A0 + nfft, B, nfft)
    u = capi.to_half_storage(x * 0.25, dtype)
    scale = (1.0 + np.arange(B, dtype=np.float32)) * 0.5
    ctx.set_host_chunk(2 * nfft * 8)
    try:
        got = ctx.fused_peaks_host(u, nfft, max_peaks=K, dtype=dtype, in_scale=scale, window=1, fs=1000.0)
    finally:
        ctx.set_host_chunk(32 << 20)
    d_in, d_sc = ctx.to_device(u), ctx.to_device(scale)
    pk, ct, g, st = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32), ctx.alloc(B * 16)
    ctx.fused_peaks(d_in, B, nfft, nfft, pk, ct, max_peaks=K, dtype=dtype, in_scale=d_sc, gmax=g, stats=st,
                    window=1, fs=1000.0)
    ctx.sync()
    ref = (pk.download(capi.MAXVALUE_DTYPE, (B, K)), ct.download(np.uint32, (B, 2)),
           g.download(capi.MAXVALUE_DTYPE, (B,)), st.download(np.float32, (B, 4)))
    assert_same_records(got, ref)
    # the scale reached the kernels: the mean |X| of channel c scales with scale[c]
    plain = ctx.fused_peaks_host(u, nfft, max_peaks=K, dtype=dtype, window=1, fs=1000.0)
    assert np.allclose(got[3][:, 0], plain[3][:, 0] * scale, rtol=1e-5)
