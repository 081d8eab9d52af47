"""GPU parity of the FFT front half through the C ABI against the FP64 oracle.

Tolerances (FP32 arithmetic, FP64-generated twiddles):
  relative L2 error of the spectrum  <= 2e-7 * log2(N) + 1e-7
  max |dX| / N                        < 1e-5   (the reference's own bound,
                                               tests/test_fft_cpu_reference_rocm.hpp:155-157)
"""
import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(99)


def crandn(*shape):
    return ((RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)) / np.sqrt(2)).astype(np.complex64)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def tol(n):
    return 2e-7 * np.log2(n) + 1e-7


@pytest.mark.parametrize("log2n", list(range(1, 19)))
def test_fft_c2c_all_sizes(ctx, log2n):
    n = 1 << log2n
    batch = max(1, min(8, (1 << 20) // n))
    print(f"[case] fft_c2c n={n} batch={batch}", flush=True)
    x = crandn(batch, n)
    d_in = ctx.to_device(x)
    d_out = ctx.alloc(x.nbytes)
    ctx.fft_c2c(d_in, d_out, batch, n, n, 0)
    ctx.sync()
    got = d_out.download(np.complex64, (batch, n))
    ref = orc.spectrum64(x, n, 0)
    assert rel_l2(got, ref) < tol(n)
    assert np.max(np.abs(got - ref)) / n < 1e-5


@pytest.mark.parametrize("n_point,nfft,win", [(1000, 1024, 0), (1000, 2048, 1), (300, 512, 2), (5000, 8192, 3),
                                              (40000, 65536, 1), (70000, 131072, 0), (1, 16, 1)])
def test_fft_zero_pad_and_windows(ctx, n_point, nfft, win):
    print(f"[case] pad n_point={n_point} nfft={nfft} window={win}", flush=True)
    batch = 3
    x = crandn(batch, n_point)
    d_in = ctx.to_device(x)
    d_out = ctx.alloc(batch * nfft * 8)
    ctx.fft_c2c(d_in, d_out, batch, n_point, nfft, win)
    ctx.sync()
    got = d_out.download(np.complex64, (batch, nfft))
    ref = orc.spectrum64(x, nfft, win)
    assert rel_l2(got, ref) < tol(nfft)


@pytest.mark.parametrize("nfft", [1024, 4096, 65536])
def test_fft_magnitude_and_magphase(ctx, nfft):
    print(f"[case] magnitude/magphase nfft={nfft}", flush=True)
    batch, n_point = 4, nfft - 7
    x = crandn(batch, n_point)
    d_in = ctx.to_device(x)
    ref = orc.spectrum64(x, nfft, 1)
    d_m = ctx.alloc(batch * nfft * 4)
    ctx.fft_magnitude(d_in, d_m, batch, n_point, nfft, 1, squared=False, scale=1.0)
    ctx.sync()
    m = d_m.download(np.float32, (batch, nfft))
    assert rel_l2(m, np.abs(ref)) < tol(nfft)
    ctx.fft_magnitude(d_in, d_m, batch, n_point, nfft, 1, squared=True, scale=0.5)
    ctx.sync()
    m2 = d_m.download(np.float32, (batch, nfft))
    assert rel_l2(m2, 0.5 * np.abs(ref) ** 2) < 2 * tol(nfft)
    d_mp = ctx.alloc(batch * nfft * 8)
    ctx.fft_mag_phase(d_in, d_mp, batch, n_point, nfft, 1)
    ctx.sync()
    mp = d_mp.download(np.float32, (batch, nfft, 2))
    assert rel_l2(mp[..., 0], np.abs(ref)) < tol(nfft)
    big = np.abs(ref) > 1e-3 * np.abs(ref).max()
    dphi = np.angle(np.exp(1j * (mp[..., 1] - np.angle(ref))))
    assert np.max(np.abs(dphi[big])) < 1e-2  # reference test tolerance (test_fft_processor_rocm.hpp:167-168)


def test_fft_closed_form_tone_on_gpu(ctx):
    n, k0, amp, ph = 65536, 12345, 1.25, 0.3
    t = np.arange(n)
    x = (amp * np.exp(1j * (2 * np.pi * k0 * t / n + ph))).astype(np.complex64)[None]
    d_in = ctx.to_device(x)
    d_out = ctx.alloc(x.nbytes)
    ctx.fft_c2c(d_in, d_out, 1, n, n, 0)
    ctx.sync()
    X = d_out.download(np.complex64, (n,))
    assert abs(X[k0] - amp * n * np.exp(1j * ph)) < 1e-4 * n
    assert np.max(np.abs(np.delete(X, k0))) < 1e-4 * n


def test_fft_linearity_full_size(ctx):
    """Size-independent property at the headline shape: FFT(a x + b y) = a X + b Y."""
    batch, n = 64, 65536
    x = crandn(batch, n)
    y = crandn(batch, n)
    a, b = 0.75, -1.5
    z = (a * x + b * y).astype(np.complex64)
    outs = []
    for arr in (x, y, z):
        d_in = ctx.to_device(arr)
        d_out = ctx.alloc(arr.nbytes)
        ctx.fft_c2c(d_in, d_out, batch, n, n, 0)
        ctx.sync()
        outs.append(d_out.download(np.complex64, (batch, n)))
    X, Y, Z = outs
    assert rel_l2(Z, a * X.astype(np.complex128) + b * Y) < 4 * tol(n)
    # Parseval per channel
    e_t = np.sum(np.abs(x.astype(np.complex128)) ** 2, axis=1) * n
    e_f = np.sum(np.abs(X.astype(np.complex128)) ** 2, axis=1)
    assert np.max(np.abs(e_f - e_t) / e_t) < 1e-5


def test_elementwise_mag_phase(ctx):
    z = np.array([0, 1, -1, 1j, -1j, 3 + 4j, 1e-6, -2 - 2j, 1e30, 5], dtype=np.complex64)
    d = ctx.to_device(z)
    o = ctx.alloc(z.size * 8)
    ctx.mag_phase(d, o, z.size)
    ctx.sync()
    mp = o.download(np.float32, (z.size, 2))
    # the reference formula sqrt(re*re + im*im) in FP32 (fft_processor_kernels_rocm.hpp:133):
    # |1e30| overflows to inf there too
    with np.errstate(over="ignore"):
        ref_mag = np.sqrt(z.real * z.real + z.imag * z.imag)
    assert np.isinf(mp[8, 0]) and np.isinf(ref_mag[8])
    fin = np.isfinite(ref_mag)
    assert np.allclose(mp[fin, 0], ref_mag[fin], rtol=1e-6, atol=1e-12)
    assert np.allclose(mp[:, 1], np.angle(z), atol=1e-6)
    assert abs(mp[5, 0] - 5.0) < 1e-6
    m = ctx.alloc(z.size * 4)
    ctx.magnitude(d, m, z.size, inv_n=0.1, squared=False)
    ctx.sync()
    mg = m.download(np.float32, (z.size,))
    assert np.isinf(mg[8])  # same FP32 overflow as the reference formula
    assert np.allclose(mg[fin], 0.1 * ref_mag[fin], rtol=1e-6)
    ctx.magnitude(d, m, z.size, inv_n=1.0, squared=True)
    ctx.sync()
    assert np.allclose(m.download(np.float32, (z.size,))[:8], np.abs(z[:8].astype(np.complex128)) ** 2, rtol=1e-6)


def test_elementwise_large_odd(ctx):
    total = (1 << 22) + 3
    z = crandn(total)
    d = ctx.to_device(z)
    o = ctx.alloc(total * 8)
    ctx.mag_phase(d, o, total)
    ctx.sync()
    mp = o.download(np.float32, (total, 2))
    assert np.allclose(mp[:, 0], np.abs(z), rtol=4e-7, atol=0)  # a few ulp
    assert np.allclose(mp[:, 1], np.angle(z), atol=2e-6)


def test_errors_are_reported(ctx):
    d = ctx.alloc(64)
    with pytest.raises(ValueError):
        ctx.fft_c2c(d, d, 1, 8, 8, 0)  # in == out
    o = ctx.alloc(64)
    with pytest.raises(ValueError):
        ctx.fft_c2c(d, o, 1, 8, 12, 0)  # not a power of two
    with pytest.raises(ValueError):
        ctx.fft_c2c(d, o, 1, 9, 8, 0)  # n_point > nfft
    with pytest.raises(ValueError):
        ctx.fft_c2c(None, o, 1, 8, 8, 0)
    ctx.fft_c2c(d, o, 0, 8, 8, 0)  # empty batch is a no-op
