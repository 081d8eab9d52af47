"""GPU checks of the learned sub-bin refiner.

* the forward/backward kernel's gradient and loss equal the FP64 oracle's
  (which is itself pinned by finite differences on the CPU);
* deployed FP16 inference equals the oracle forward pass of the FP16-rounded
  weights to FP16 accuracy;
* features and labels of detected peaks equal the oracle's, and the labels
  agree with the generator's true tone frequencies;
* a network trained on the GPU beats parabolic refinement on held-out peaks,
  and the fused peak kernel with refine = LEARNED reproduces its predictions.
"""
import numpy as np
import pytest

from spectrum_amd import capi
from oracle import mlp as omlp
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_oracle(ctx):
    from spectrum_amd.refiner import Refiner

    rng = np.random.default_rng(3)
    r = Refiner(ctx, seed=11)
    p = r.get_params().astype(np.float64)
    p[omlp.OFF_B1:omlp.OFF_W2] = rng.normal(0, 0.1, omlp.H)  # non-zero biases exercise every term
    p[omlp.OFF_B2:omlp.OFF_W3] = rng.normal(0, 0.1, omlp.H)
    p[omlp.OFF_B3] = 0.05
    r.set_params(p.astype(np.float32))
    p32 = r.get_params().astype(np.float64)
    n = 3000
    X = rng.normal(0, 0.7, (n, 10)).astype(np.float32)
    Y = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    dX, dY = ctx.to_device(X), ctx.to_device(Y)
    g = ctx.alloc((omlp.NPARAM + 1) * 4)
    ctx.mlp_grad(r.params, dX, dY, n, g)
    ctx.sync()
    got = g.download(np.float32, (omlp.NPARAM + 1,))
    ref = omlp.grad(p32, X.astype(np.float64), Y.astype(np.float64))
    ref_loss = omlp.loss(p32, X.astype(np.float64), Y.astype(np.float64))
    assert abs(got[-1] - ref_loss) <= 1e-5 * ref_loss
    scale = np.abs(ref).max()
    assert np.max(np.abs(got[:-1] - ref)) <= 2e-5 * scale, np.max(np.abs(got[:-1] - ref)) / scale


def test_mlp_deployed_inference_matches_oracle(ctx):
    from spectrum_amd.refiner import Refiner

    rng = np.random.default_rng(4)
    r = Refiner(ctx, seed=12)
    r.export()
    X = rng.normal(0, 0.5, (1000, 10)).astype(np.float32)
    got = r.predict(X)
    p = r.get_params().astype(np.float16).astype(np.float64)
    Xh = X.astype(np.float16).astype(np.float64)
    ref, _ = omlp.forward(p, Xh)
    err = np.abs(got - ref)
    assert np.max(err) <= 5e-3 * (1.0 + np.abs(ref).max()), np.max(err)


def test_dataset_features_and_labels(ctx):
    from spectrum_amd.refiner import make_peaks

    N, K, B = 8192, 64, 6
    ps = make_peaks(ctx, B, 0x5EED0105, N=N, K=K, chunk=B)
    assert len(ps) > 10
    x = capi.siggen(capi.GEN_FRACTIONAL, 0x5EED0105, B, N)
    X64 = orc.spectrum64(x, N, 1)
    w = orc.window(1, N)
    checked = 0
    for i in range(0, len(ps), max(1, len(ps) // 12)):
        c, k = int(ps.channel[i]), int(ps.bin[i])
        Xc = X64[c]
        assert np.allclose(omlp.features(Xc, k), ps.feat[i], atol=2e-4)
        xp = np.zeros(16 * N, np.complex128)
        xp[:N] = x[c] * w
        lab = omlp.label(np.abs(orc.fft64(xp)), k)
        assert lab is not None and abs(lab - ps.label[i]) < 2e-3
        checked += 1
    assert checked >= 5
    ok = np.isfinite(ps.truth)
    assert np.sqrt(np.mean((ps.label[ok] - ps.truth[ok]) ** 2)) < 0.01


def test_trained_refiner_beats_parabolic_and_fused_inference(ctx):
    from spectrum_amd.refiner import Refiner, make_peaks, rmse

    N, K = 8192, 64
    train = make_peaks(ctx, 6144, 0x5EED1005, N=N, K=K)
    test = make_peaks(ctx, 512, 0x5EED0005, N=N, K=K)
    print(f"[gpu] refiner data: {len(train)} training peaks, {len(test)} held-out", flush=True)
    r = Refiner(ctx, seed=5)
    losses = r.train(train.feat, train.label, steps=4000, batch=1024, lr=2e-3)
    assert losses[-200:].mean() < 0.5 * losses[:50].mean()
    r.export()
    pred = r.predict(test.feat)
    e_learn, e_par = rmse(pred, test.label), rmse(test.parab, test.label)
    print(f"[gpu] held-out RMSE vs 16x padded label: learned {e_learn:.4f} bin, parabolic {e_par:.4f} bin",
          flush=True)
    assert e_learn < 0.7 * e_par
    assert e_learn < 0.03
    # fused kernel with the learned refiner: its records carry the same predictions
    B = 512
    x = capi.siggen(capi.GEN_FRACTIONAL, 0x5EED0005, B, N)
    d_in = ctx.to_device(x)
    pk, ct = ctx.alloc(B * K * 32), ctx.alloc(B * 8)
    ctx.set_refiner(r.deployed)
    ctx.fused_peaks(d_in, B, N, N, pk, ct, window=1, nsigma=6.0, max_peaks=K, refine=capi.REFINE_LEARNED)
    ctx.sync()
    ctx.set_refiner(None)
    recs = pk.download(capi.MAXVALUE_DTYPE, (B, K))
    got = recs["freq_offset"][test.channel, test.slot]
    use = np.abs(pred) <= 0.5
    assert np.max(np.abs(got[use] - pred[use])) < 2e-3
    assert np.allclose(recs["refined_frequency"][test.channel, test.slot],
                       (recs["index"][test.channel, test.slot] + got) / N, rtol=1e-5, atol=1e-7)


def test_fused_learned_many_peaks_matches_predict(ctx):
    """More than 64 stored peaks per channel (nsigma = 0 on tones in noise):
    the records equal the parabolic ones except the offset, which is the
    network's prediction on the peak's features wherever that lies in
    [-0.5, 0.5] and the parabolic offset elsewhere."""
    from spectrum_amd.refiner import DEFAULT_WEIGHTS, Refiner

    N, K, B = 4096, 100, 4
    print(f"[case] learned many peaks N={N} K={K}", flush=True)
    rng = np.random.default_rng(106)
    n = np.arange(N)
    x = (rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))) * 0.3
    for f in (0.1031, 0.2277, -0.3519):
        x = x + 4.0 * np.exp(2j * np.pi * f * n)
    x = x.astype(np.complex64)
    r = Refiner(ctx).load(DEFAULT_WEIGHTS)
    r.export()
    d_in = ctx.to_device(x)
    out = {}
    for name, refine in (("learned", capi.REFINE_LEARNED), ("parabolic", capi.REFINE_PARABOLIC)):
        pk, ct = ctx.alloc(B * K * 32), ctx.alloc(B * 8)
        ctx.set_refiner(r.deployed if refine == capi.REFINE_LEARNED else None)
        ctx.fused_peaks(d_in, B, N, N, pk, ct, window=1, nsigma=0.0, max_peaks=K, refine=refine)
        ctx.sync()
        out[name] = (pk.download(capi.MAXVALUE_DTYPE, (B, K)), ct.download(np.uint32, (B, 2)))
    ctx.set_refiner(None)
    (L, cl), (P, cp) = out["learned"], out["parabolic"]
    assert np.array_equal(cl, cp) and np.all(cl[:, 0] == K)
    for f in ("index", "real", "imag", "magnitude", "phase"):
        assert L[f].tobytes() == P[f].tobytes(), f
    spec = ctx.alloc(B * N * 8)
    ctx.fft_c2c(d_in, spec, B, N, N, 1)
    ctx.sync()
    X = spec.download(np.complex64, (B, N)).astype(np.complex128)
    feats = np.array([omlp.features(X[c], int(k)) for c in range(B) for k in L["index"][c]])
    pred = r.predict(feats.astype(np.float32)).reshape(B, K)
    got = L["freq_offset"]
    inside = np.abs(pred) <= 0.5 - 2e-3
    outside = np.abs(pred) > 0.5 + 2e-3
    assert inside[:, 64:].any() and inside.sum() >= B * 16  # rows past 64 are covered
    assert np.max(np.abs(got[inside] - pred[inside])) < 2e-3
    assert got[outside].tobytes() == P["freq_offset"][outside].tobytes()
    assert np.allclose(L["refined_frequency"], (L["index"] + got) / N, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("N", [1 << 15, 1 << 16, 1 << 17, 1 << 18, 1 << 19, 1 << 20])
def test_split_learned_matches_predict(ctx, N):
    """Learned refinement on the split long-channel path (features from the
    row kernel's candidate neighbourhoods and complex block-edge rows, the
    refiner once per chunk) and, from 2^19, on the Stockham path (features
    from the stored chunk spectrum in the gather kernel).  Records equal the parabolic ones except the
    offset, which is the network's prediction on the features of the
    spectrum (the split dense transform, same arithmetic) wherever it lies in
    [-0.5, 0.5].  Two regimes: tones on row-block edges (k2 = k mod N2 in
    {0, 1, 14, 15} mod 16 and bins 0, 1, N-2, N-1: neighbours in other blocks),
    threshold mean + 6 std (few candidates); and nsigma = 0 on noise, more
    than 64 candidates (the bitmap selection)."""
    from spectrum_amd.refiner import DEFAULT_WEIGHTS, Refiner

    N2 = N // 256
    r = Refiner(ctx).load(DEFAULT_WEIGHTS)
    r.export()
    n = np.arange(N)
    rng = np.random.default_rng(N + 7)
    bins = [[5 * N2 + 0, 9 * N2 + 1, 200 * N2 + 14], [3 * N2 + 15, 0, N // 2 + 7], [N - 1, 60 * N2 + 1, 120 * N2 + 15],
            [1, 31 * N2 + 17, 250 * N2 + 30], [N - 2, 100 * N2 + 31, 150 * N2 + 16]]
    B = len(bins)
    x = 0.01 * (rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N)))
    for b, ks in enumerate(bins):
        for j, k in enumerate(ks):
            x[b] += (1.0 - 0.2 * j) * np.exp(2j * np.pi * (k + 0.1 + 0.15 * j) * n / N)
    noise = (rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))).astype(np.complex64)
    for name, sig, nsigma, K in (("edges", x.astype(np.complex64), 6.0, 64), ("noise", noise, 0.0, 100)):
        Bc = sig.shape[0]
        print(f"[case] split learned N={N} {name}", flush=True)
        d_in = ctx.to_device(sig)
        out = {}
        for rname, refine in (("learned", capi.REFINE_LEARNED), ("parabolic", capi.REFINE_PARABOLIC)):
            pk, ct = ctx.alloc(Bc * K * 32), ctx.alloc(Bc * 8)
            ctx.set_refiner(r.deployed if refine == capi.REFINE_LEARNED else None)
            ctx.fused_peaks(d_in, Bc, N, N, pk, ct, window=1, nsigma=nsigma, max_peaks=K, refine=refine)
            ctx.sync()
            out[rname] = (pk.download(capi.MAXVALUE_DTYPE, (Bc, K)), ct.download(np.uint32, (Bc, 2)))
        ctx.set_refiner(None)
        (L, cl), (P, cp) = out["learned"], out["parabolic"]
        assert np.array_equal(cl, cp)
        if name == "noise":
            assert np.all(cl[:, 0] == K)
        else:
            for b, ks in enumerate(bins):
                assert set(ks) <= set(int(k) for k in L["index"][b, : cl[b, 0]]), (b, ks)
        for b in range(Bc):
            for f in ("index", "real", "imag", "magnitude", "phase"):
                assert L[f][b, : cl[b, 0]].tobytes() == P[f][b, : cl[b, 0]].tobytes(), (b, f)
        spec = ctx.alloc(Bc * N * 8)
        ctx.fft_c2c(d_in, spec, Bc, N, N, 1)
        ctx.sync()
        X = spec.download(np.complex64, (Bc, N)).astype(np.complex128)
        rows = [(c, s) for c in range(Bc) for s in range(int(cl[c, 0]))]
        feats = np.array([omlp.features(X[c], int(L["index"][c, s])) for c, s in rows])
        pred = r.predict(feats.astype(np.float32))
        got = np.array([L["freq_offset"][c, s] for c, s in rows])
        par = np.array([P["freq_offset"][c, s] for c, s in rows])
        inside = np.abs(pred) <= 0.5 - 2e-3
        outside = np.abs(pred) > 0.5 + 2e-3
        assert inside.sum() >= len(rows) // 2
        assert np.max(np.abs(got[inside] - pred[inside])) < 2e-3
        assert np.array_equal(got[outside], par[outside])
        for b in range(Bc):
            m = int(cl[b, 0])
            assert np.allclose(L["refined_frequency"][b, :m], (L["index"][b, :m] + L["freq_offset"][b, :m]) / N,
                               rtol=1e-5, atol=1e-7)
