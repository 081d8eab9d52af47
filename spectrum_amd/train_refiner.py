"""Train the learned sub-bin refiner on the GPU and evaluate it.

    python -m spectrum_amd.train_refiner --out spectrum_amd/data/refiner_hann_8192.npz \
        [--peaks 1048576] [--steps 20000] [--report report.json]

Training data: channels of 8192 samples from the library generator (1-16
tones at fractional bins, 20 dB above the noise, Hann window), peaks found by
the fused threshold search (mean + 6 std), labels from the 16x zero-padded
transform.  Evaluation on 4096 held-out channels (a different seed): RMSE of
the learned and the parabolic offsets against the padded-FFT label and
against the generator's true frequency, plus the zero-padding factor that
parabolic refinement needs to reach the learned RMSE and the fused kernel
time of each form.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import capi
from .refiner import Refiner, make_peaks, rmse

N, K = 8192, 64
TRAIN_SEED, TEST_SEED = 0x5EED1005, 0x5EED0005


def _time_fused(ctx, d_in, B, n_point, nfft, pk, ct, refine, reps=10):
    hip = __import__("ctypes").CDLL("libamdhip64.so")
    import ctypes as C

    ev = [C.c_void_p(), C.c_void_p()]
    for e in ev:
        hip.hipEventCreate(C.byref(e))
    st = C.c_void_p(ctx.stream())
    for _ in range(2):
        ctx.fused_peaks(d_in, B, n_point, nfft, pk, ct, window=1, nsigma=6.0, max_peaks=K, refine=refine)
    ctx.sync()
    ts = []
    for _ in range(reps):
        hip.hipEventRecord(ev[0], st)
        ctx.fused_peaks(d_in, B, n_point, nfft, pk, ct, window=1, nsigma=6.0, max_peaks=K, refine=refine)
        hip.hipEventRecord(ev[1], st)
        hip.hipEventSynchronize(ev[1])
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), ev[0], ev[1])
        ts.append(ms.value)
    return float(np.median(ts))


def evaluate(ctx, r: Refiner, n_chan: int = 4096) -> dict:
    test = make_peaks(ctx, n_chan, TEST_SEED, N=N, K=K)
    r.export()
    pred = r.predict(test.feat)
    ok = np.isfinite(test.truth)
    out = {
        "held_out_channels": n_chan, "held_out_peaks": int(len(test)),
        "rmse_learned_vs_label": rmse(pred, test.label), "rmse_parabolic_vs_label": rmse(test.parab, test.label),
        "rmse_learned_vs_truth": rmse(pred[ok], test.truth[ok]),
        "rmse_parabolic_vs_truth": rmse(test.parab[ok], test.truth[ok]),
        "rmse_label_vs_truth": rmse(test.label[ok], test.truth[ok]),
    }
    # fused kernel: learned at nfft = N against parabolic with zero padding P
    x, tf, ta = capi.siggen(capi.GEN_FRACTIONAL, TEST_SEED, n_chan, N, max_tones=16, with_tones=True)
    d_in = ctx.to_device(x)
    pk, ct = ctx.alloc(n_chan * K * 32), ctx.alloc(n_chan * 8)
    ctx.set_refiner(r.deployed)
    t_learn = _time_fused(ctx, d_in, n_chan, N, N, pk, ct, capi.REFINE_LEARNED)
    ctx.set_refiner(None)
    rows = [{"refine": "learned", "zero_pad": 1, "ms": t_learn,
             "rmse_vs_truth": _rmse_truth(ctx, d_in, pk, ct, n_chan, N, tf, ta, 1, capi.REFINE_LEARNED, r)}]
    for P in (1, 2, 4, 8, 16):
        t = _time_fused(ctx, d_in, n_chan, N, N * P, pk, ct, capi.REFINE_PARABOLIC)
        rows.append({"refine": "parabolic", "zero_pad": P, "ms": t,
                     "rmse_vs_truth": _rmse_truth(ctx, d_in, pk, ct, n_chan, N, tf, ta, P, capi.REFINE_PARABOLIC,
                                                  None)})
    out["fused"] = rows
    need = [row["zero_pad"] for row in rows[1:] if row["rmse_vs_truth"] <= rows[0]["rmse_vs_truth"]]
    out["parabolic_zero_pad_for_equal_rmse"] = need[0] if need else None
    return out


def _rmse_truth(ctx, d_in, pk, ct, B, n_point, tf, ta, P, refine, r):
    """Refined frequency of every stored peak against the nearest true tone (within one coarse bin)."""
    if refine == capi.REFINE_LEARNED:
        ctx.set_refiner(r.deployed)
    ctx.fused_peaks(d_in, B, n_point, n_point * P, pk, ct, window=1, nsigma=6.0, max_peaks=K, refine=refine)
    ctx.sync()
    ctx.set_refiner(None)
    recs = pk.download(capi.MAXVALUE_DTYPE, (B, K))
    cnt = ct.download(np.uint32, (B, 2))[:, 0]
    errs = []
    for b in range(B):
        f = recs["refined_frequency"][b, : cnt[b]].astype(np.float64) * n_point  # coarse bins (fs = 1)
        tb = np.mod(tf[b][ta[b] > 0] * n_point, n_point)
        if len(f) == 0 or len(tb) == 0:
            continue
        d = (f[:, None] - tb[None, :] + n_point / 2) % n_point - n_point / 2
        j = np.argmin(np.abs(d), axis=1)
        e = d[np.arange(len(j)), j]
        errs.append(e[np.abs(e) <= 1.0])
    e = np.concatenate(errs) if errs else np.array([np.nan])
    return float(np.sqrt(np.mean(e * e)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--peaks", type=int, default=1 << 20)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    ctx = capi.Context(0)
    t0 = time.perf_counter()
    chunks, n = [], 0
    c0 = 0
    while n < args.peaks:
        ps = make_peaks(ctx, 8192, TRAIN_SEED, N=N, K=K, chan_offset=c0)
        chunks.append(ps)
        n += len(ps)
        c0 += 8192
        print(f"[train] {n} peaks from {c0} channels", flush=True)
    feat = np.concatenate([p.feat for p in chunks])[: args.peaks]
    label = np.concatenate([p.label for p in chunks])[: args.peaks]
    t_data = time.perf_counter() - t0
    r = Refiner(ctx, seed=5)
    t1 = time.perf_counter()
    losses = []
    for i in range(0, args.steps, 2000):
        s = min(2000, args.steps - i)
        lr = args.lr * (0.1 if i >= 0.8 * args.steps else 1.0)
        losses.append(r.train(feat, label, steps=s, batch=args.batch, lr=lr))
        print(f"[train] step {i + s}: loss {losses[-1][-100:].mean():.3e}", flush=True)
    t_train = time.perf_counter() - t1
    rep = {"train_peaks": int(len(label)), "train_channels": c0, "steps": args.steps, "batch": args.batch,
           "data_s": t_data, "train_s": t_train, "final_loss": float(np.concatenate(losses)[-500:].mean())}
    rep.update(evaluate(ctx, r))
    r.save(args.out, train_peaks=len(label), steps=args.steps)
    print(json.dumps(rep, indent=1), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()
