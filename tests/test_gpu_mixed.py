"""GPU checks of the mixed-precision path (fp16/bf16 storage, DFT-as-GEMM on MFMA), N = 16384.

Spectral error is the relative L2 error ||X - X64|| / ||X64|| per channel,
where X64 is the FP64 transform of the same windowed input: "transform" uses
the quantized input the kernel actually reads, "end-to-end" the complex64
signal before quantization.  Bound: 1e-3 for the transform error with either
storage and either compute type; end-to-end adds the storage rounding itself
(about 2e-4 for fp16, 1.3e-3 for bf16).  The stress signal is one strong tone,
one tone 60 dB weaker and noise 80 dB below the strong tone.
"""
import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

N = 16384


def rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


@pytest.mark.parametrize("dtype", [capi.DTYPE_C32F16, capi.DTYPE_C32BF16])
@pytest.mark.parametrize("compute", [capi.COMPUTE_F16, capi.COMPUTE_BF16])
def test_mixed_spectrum_error(ctx, dtype, compute):
    B = 8
    x = capi.siggen(capi.GEN_DYNAMIC, 0x5EED0004, B, N)
    u = capi.to_half_storage(x, dtype)
    xq = capi.from_half_storage(u, dtype)
    d_in = ctx.to_device(u)
    d_out = ctx.alloc(B * N * 8)
    print(f"[gpu] fft_mixed dtype={dtype} compute={compute}", flush=True)
    ctx.fft_mixed(d_in, d_out, B, N, window=1, dtype=dtype, compute=compute)
    ctx.sync()
    X = d_out.download(np.complex64, (B, N))
    ref_q = orc.spectrum64(xq, N, 1)
    ref_x = orc.spectrum64(x, N, 1)
    e_t = rel_l2(X, ref_q)
    e_e = rel_l2(X, ref_x)
    print(f"[gpu] transform err max {e_t.max():.2e}, end-to-end max {e_e.max():.2e}", flush=True)
    assert np.all(e_t <= 1e-3), e_t
    assert np.all(e_e <= (1e-3 if dtype == capi.DTYPE_C32F16 else 3e-3)), e_e
    # the FP32 path on the same stored input is far more accurate; the MFMA form must still find both tones
    for b in range(B):
        m = np.abs(X[b])
        mq = np.abs(ref_q[b])
        top = int(np.argmax(mq))
        assert int(np.argmax(m)) == top


def test_mixed_closed_forms(ctx):
    """Impulse -> flat spectrum; integer-bin tone -> A*N at its bin (no window)."""
    B = 2
    x = np.zeros((B, N), np.complex64)
    x[0, 0] = 1.0
    k0 = 1234
    x[1] = np.exp(2j * np.pi * k0 * np.arange(N) / N).astype(np.complex64) * 0.5
    u = capi.to_half_storage(x, capi.DTYPE_C32F16)
    d_in = ctx.to_device(u)
    d_out = ctx.alloc(B * N * 8)
    ctx.fft_mixed(d_in, d_out, B, N, window=0, dtype=capi.DTYPE_C32F16)
    ctx.sync()
    X = d_out.download(np.complex64, (B, N))
    assert np.max(np.abs(X[0] - 1.0)) < 2e-3
    assert abs(X[1][k0] - 0.5 * N) < 1e-3 * N
    other = np.delete(np.abs(X[1]), k0)
    assert other.max() < 2e-3 * N


@pytest.mark.parametrize("dtype", [capi.DTYPE_C32F16, capi.DTYPE_C32BF16])
def test_mixed_peaks_vs_oracle_and_fp32(ctx, dtype):
    """Peak lists of the MFMA path against the oracle's rules on the FP64 spectrum of the stored input
    (tie tolerance widened to the FP16 transform error) and against the FP32 kernel on the same storage."""
    B, K = 16, 32
    x = capi.siggen(capi.GEN_DYNAMIC, 0x5EED0014, B, N)
    u = capi.to_half_storage(x, dtype)
    xq = capi.from_half_storage(u, dtype)
    d_in = ctx.to_device(u)
    outs = {}
    for name in ("mixed", "fp32"):
        pk, ct, gm, st = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32), ctx.alloc(B * 16)
        if name == "mixed":
            ctx.fused_peaks_mixed(d_in, B, N, pk, ct, dtype=dtype, nsigma=6.0, max_peaks=K, gmax=gm, stats=st)
        else:
            ctx.fused_peaks(d_in, B, N, N, pk, ct, dtype=dtype, nsigma=6.0, max_peaks=K, gmax=gm, stats=st)
        ctx.sync()
        outs[name] = (pk.download(capi.MAXVALUE_DTYPE, (B, K)), ct.download(np.uint32, (B, 2)),
                      gm.download(capi.MAXVALUE_DTYPE, (B,)), st.download(np.float32, (B, 4)))
    X = orc.spectrum64(xq, N, 1)
    agree_oracle = agree_fp32 = 0
    for b in range(B):
        recs, found, g, stats, mags = orc.fused_peaks(X[b], 1.0, nsigma=6.0, kcap=K)
        tau = 2e-3 * mags.max()
        pm, cm, gmm, sm = (o[b] for o in outs["mixed"])
        pf, cf, gf, sf = (o[b] for o in outs["fp32"])
        got = pm[: cm[0]]["index"]
        ok, why = orc.compare_peak_lists(got, recs["index"], mags, stats[2], tau, K)
        assert ok, f"channel {b}: {why}"
        agree_oracle += int(list(got) == list(recs["index"]))
        agree_fp32 += int(list(got) == list(pf[: cf[0]]["index"]))
        assert gmm["index"] == g["index"]
        assert abs(sm[2] - stats[2]) <= 2e-3 * stats[2]
        assert abs(int(cm[1]) - found) <= 1
    print(f"[gpu] exact peak-list agreement: oracle {agree_oracle}/{B}, fp32 path {agree_fp32}/{B}", flush=True)


def test_mixed_rejects_unsupported(ctx):
    d = ctx.alloc(16)
    with pytest.raises(capi.SpxError):
        ctx.fft_mixed(d, d, 1, 8192, dtype=capi.DTYPE_C32F16)
    with pytest.raises(ValueError):
        ctx.fft_mixed(d, d, 1, N, dtype=capi.DTYPE_C64)
