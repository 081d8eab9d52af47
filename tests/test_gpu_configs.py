"""The mixed-precision and learned-refinement workloads checked on their full
batches (not samples):

* 2048 x 16384 stress signal (one strong tone, one 60 dB weaker, noise 80 dB
  below), fp16 and bf16 storage, MFMA form: every channel's peak list against
  the oracle's rules on the FP64 spectrum of the stored input (tie tolerance
  widened to the FP16 transform error, 2e-3 max|X|), the transform's relative
  spectral L2 error <= 1e-3 (64 channels), and learned refinement within 0.02
  bin RMS of the generator's true tones and better than parabolic;
* 4096 x 8192 fractional-bin tones (the learned-refinement workload), FP32:
  every channel against the oracle, then learned refinement over the whole
  batch: records equal the parabolic ones except the offset, RMSE against
  the true tones well below parabolic's.
"""
import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc
from test_gpu_peaks import check_channels, run_fused

pytestmark = pytest.mark.gpu


def _rmse(recs, cnt, tf, ta, N):
    errs = []
    for b in range(recs.shape[0]):
        f = recs["refined_frequency"][b, : cnt[b]].astype(np.float64) * N
        tb = np.mod(tf[b][ta[b] > 0] * N, N)
        if len(f) and len(tb):
            dd = (f[:, None] - tb[None, :] + N / 2) % N - N / 2
            e = dd[np.arange(len(f)), np.argmin(np.abs(dd), axis=1)]
            errs.append(e[np.abs(e) <= 1.0])
    e = np.concatenate(errs)
    return float(np.sqrt(np.mean(e * e))), len(e)


@pytest.mark.parametrize("dtype", [capi.DTYPE_C32F16, capi.DTYPE_C32BF16])
def test_config4_full_batch(ctx, dtype):
    from spectrum_amd.refiner import DEFAULT_WEIGHTS, Refiner

    B, N, K = 2048, 16384, 32
    print(f"[case] config4 full batch dtype={dtype}", flush=True)
    x, tf, ta = capi.siggen(capi.GEN_DYNAMIC, 0x5EED0004, B, N, max_tones=2, with_tones=True)
    u = capi.to_half_storage(x, dtype)
    xq = capi.from_half_storage(u, dtype)
    d16 = ctx.to_device(u)
    r = Refiner(ctx).load(DEFAULT_WEIGHTS)
    r.export()
    res = {}
    for name, refine in (("parabolic", capi.REFINE_PARABOLIC), ("learned", capi.REFINE_LEARNED)):
        pk, ct = ctx.alloc(B * K * 32), ctx.alloc(B * 8)
        ctx.set_refiner(r.deployed if refine == capi.REFINE_LEARNED else None)
        ctx.fused_peaks_mixed(d16, B, N, pk, ct, dtype=dtype, max_peaks=K, window=1, refine=refine)
        ctx.sync()
        res[name] = (pk.download(capi.MAXVALUE_DTYPE, (B, K)), ct.download(np.uint32, (B, 2)))
    ctx.set_refiner(None)
    P, C = res["parabolic"]
    X = orc.spectrum64(xq, N, 1)
    for b in range(B):
        recs, found, g, stats, mags = orc.fused_peaks(X[b], 1.0, nsigma=6.0, kcap=K)
        ok, why = orc.compare_peak_lists(P[b, : C[b, 0]]["index"], recs["index"], mags, stats[2], 2e-3 * mags.max(),
                                         K)
        assert ok, f"channel {b}: {why}"
    S = 64
    sp = ctx.alloc(S * N * 8)
    ctx.fft_mixed(d16, sp, S, N, window=1, dtype=dtype)
    ctx.sync()
    Xg = sp.download(np.complex64, (S, N))
    e_t = np.linalg.norm(Xg - X[:S], axis=1) / np.linalg.norm(X[:S], axis=1)
    assert np.all(e_t <= 1e-3), e_t.max()
    L, CL = res["learned"]
    assert np.array_equal(C, CL)
    e_learn, n_l = _rmse(L, CL[:, 0], tf, ta, N)
    e_par, _ = _rmse(P, C[:, 0], tf, ta, N)
    print(f"[gpu] C4 dtype={dtype}: RMSE learned {e_learn:.5f} bin, parabolic {e_par:.5f} bin over {n_l} peaks",
          flush=True)
    assert n_l >= B
    assert e_learn <= 0.02 and e_learn < 0.5 * e_par


def test_config5_full_batch(ctx):
    from spectrum_amd.refiner import DEFAULT_WEIGHTS, Refiner

    B, N, K = 4096, 8192, 64
    print("[case] config5 full batch", flush=True)
    x, tf, ta = capi.siggen(capi.GEN_FRACTIONAL, 0x5EED0005, B, N, max_tones=16, with_tones=True)
    P, C, G, S = run_fused(ctx, x, N, K=K)
    check_channels(x, N, P, C, G, S, K=K, chunk=256)
    r = Refiner(ctx).load(DEFAULT_WEIGHTS)
    r.export()
    ctx.set_refiner(r.deployed)
    L, CL, GL, SL = run_fused(ctx, x, N, K=K, refine=capi.REFINE_LEARNED)
    ctx.set_refiner(None)
    assert np.array_equal(C, CL)
    for b in range(B):
        n = int(C[b, 0])
        for f in ("index", "real", "imag", "magnitude", "phase"):
            assert L[f][b, :n].tobytes() == P[f][b, :n].tobytes(), (b, f)
    e_learn, n_l = _rmse(L, CL[:, 0], tf, ta, N)
    e_par, _ = _rmse(P, C[:, 0], tf, ta, N)
    print(f"[gpu] C5: RMSE learned {e_learn:.5f} bin, parabolic {e_par:.5f} bin over {n_l} peaks", flush=True)
    assert e_learn <= 0.02 and e_learn < 0.2 * e_par
