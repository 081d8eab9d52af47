"""Headline benchmark: fused window -> FFT -> |X| -> peak search on one MI355X.

Workload (the metric's configuration): 1024 channels x 65536-point complex64,
1-8 tones per channel at 0-20 dB SNR, Hann window, threshold mean + 6 std of
|X|, first 64 local maxima in ascending bin order, parabolic refinement,
global-maximum record.  One step = one call of the library over the whole
batch (three launches: column transforms, row transforms with |X| and local
maxima, per-channel finish), input already resident in HBM.

Prints one JSON line: value = channels/s, roofline = algorithmic bytes of the
call / its average time (hipEvents on the launch stream) vs 8 TB/s, with the
HBM traffic of the three kernels from the committed rocprofv3 counter
summary, cpu_baseline = the reference's CPU path (its
pocketfft CpuFFT, compiled from the reference sources into oracle/_ref) plus
the oracle's C peak rules, timed on a bounded sample on this host (one
thread, plus a multi-core variant).  Extra objects on the same line (skipped
with --no-extras):
  small_batch  64 x 4096, three tones (the CPU-runnable shape): time per
             call and the reference CPU path on the same 64 channels;
  uneven     4096 channels of lengths 2^10..2^17 in one hipGraph launch vs
             per-length launches (autotuned plan and the default one);
  mixed_mlp  2048 x 16384: the FP16-MFMA + learned-refinement form against
             the plain FP32 form (complex64, parabolic), with its spectral
             error against FP64;
  refiner    4096 x 8192 fractional tones: learned vs parabolic refinement
             (time per launch, frequency RMSE against the true tones);
  refiner_split  the same at the headline shape, 1024 x 65536 (split path);
  top_peaks  the reference ONE/TWO_PEAKS search inside the transform vs
             transform-then-search, 1024 x 65536;
  large      64 x 2^20 (past the split path): peak search and complex
             spectrum, time and fraction of 8 TB/s on algorithmic bytes.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FFT+peak-search channels/s at 1024x64K c64; %HBM peak; MFMA+MLP vs FP32"
HBM_PEAK_GBS = 8000.0


class HipTimer:
    """hipEvent pair on a given stream (the stream the kernels are launched on)."""

    def __init__(self, stream):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]
        self.stream = C.c_void_p(stream)
        self.a, self.b = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(self.a)) == 0
        assert self.hip.hipEventCreate(C.byref(self.b)) == 0

    def start(self):
        assert self.hip.hipEventRecord(self.a, self.stream) == 0

    def stop_ms(self):
        assert self.hip.hipEventRecord(self.b, self.stream) == 0
        assert self.hip.hipEventSynchronize(self.b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), self.a, self.b) == 0
        return float(ms.value)


def cpu_model() -> str:
    """Host CPU model name (from /proc/cpuinfo) and logical CPU count."""
    name = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{name}, {os.cpu_count()} logical cpus"


def _median_runs(fn, runs=5):
    """One warm-up run, then the median wall time of `runs` runs (seconds)."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(x, nfft, K, target_s=2.0, runs=5):
    """Reference CPU FFT (pocketfft float, one thread, plan per call) + oracle
    peak rules on a bounded sample of channels: 1 warm-up run, median of
    `runs` runs, the sample sized so that one run takes about target_s."""
    from oracle import oracle as orc

    w = orc.window(1, nfft).astype(np.float32)
    kind = "reference" if orc.have_reference() else "port"

    def one(b):
        xw = (x[b] * w).astype(np.complex64)[None]
        X = orc.ref_cpufft(xw)[0] if kind == "reference" else orc.spectrum64(xw, nfft, 0)[0].astype(np.complex64)
        orc.fused_peaks(X, 1.0, nsigma=6.0, kcap=K)

    t0 = time.perf_counter()
    one(0)
    per = max(time.perf_counter() - t0, 1e-6)
    S = int(min(x.shape[0], max(8, target_s / per)))
    med, ts = _median_runs(lambda: [one(b) for b in range(S)], runs)
    return {
        "value": S / med,
        "unit": "channels/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{S} of {x.shape[0]} channels x {nfft} c64 (Hann, pocketfft c2c<float> via the reference "
                  f"CpuFFT::ForwardC2C, then |X|/mean+6std/local-max/K={K}/parabolic), single thread; "
                  f"1 warm-up run, median of {runs} runs ({', '.join(f'{t:.2f}' for t in ts)} s)",
        "host": cpu_model(),
    }


def cpu_baseline_mt(x, nfft, K, threads, runs=5):
    """Best-effort multi-core CPU path: pocketfft c2c over [B, N] with `threads` threads
    (reference sources, oracle/_ref) + the oracle's peak rules over channels on a thread pool;
    1 warm-up run, median of `runs` runs over a sample of up to 512 channels."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as orc

    if not orc.have_reference():
        return None
    w = orc.window(1, nfft).astype(np.float32)
    S = min(x.shape[0], 512)
    xw = (x[:S] * w).astype(np.complex64)
    with ThreadPoolExecutor(threads) as pool:
        def one():
            X = orc.ref_pocketfft_mt(xw, threads)
            list(pool.map(lambda r: orc.fused_peaks(r, 1.0, nsigma=6.0, kcap=K), list(X)))

        med, ts = _median_runs(one, runs)
    return {"value": S / med, "unit": "channels/s", "cores": threads, "kind": "reference",
            "sample": f"{S} channels x {nfft}: pocketfft c2c<float> with {threads} threads (reference sources), "
                      f"then the peak rules on a {threads}-thread pool; 1 warm-up run, median of {runs} runs "
                      f"({', '.join(f'{t:.2f}' for t in ts)} s)"}


def _timer_ms(ctx, fn, reps=30, warm=20):
    t = HipTimer(ctx.stream())
    for _ in range(warm):
        fn()
    ctx.sync()
    ts = []
    for _ in range(reps):
        t.start()
        fn()
        ts.append(t.stop_ms())
    return float(np.median(ts))


def extra_host(ctx, x, K=64, reps=5):
    """PCIe-inclusive wall time of the headline call with the input in host
    memory and the results returned to host memory (reported beside the
    device-resident value, never as it): the streamed host entry point
    (spx_fused_peaks_host) against the plain sequence upload -> search ->
    download; the same for the complex spectrum (spx_fft_c2c_host); the pinned
    DMA rate bounds both."""
    B, N = x.shape

    def wall(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    d = ctx.alloc(x.nbytes)
    pk, ct, g = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32)

    def plain_peaks():
        d.upload(x)
        ctx.fused_peaks(d, B, N, N, pk, ct, window=1, fs=1.0, nsigma=6.0, max_peaks=K, gmax=g)
        ctx.sync()
        pk.download(np.uint8, (B * K * 32,))
        ct.download(np.uint32, (B, 2))
        g.download(np.uint8, (B * 32,))

    def streamed_peaks(src):
        ctx.fused_peaks_host(src, N, window=1, fs=1.0, nsigma=6.0, max_peaks=K, with_stats=False)

    out = {"plain_ms": wall(plain_peaks), "streamed_ms": wall(lambda: streamed_peaks(x))}
    xp = ctx.pinned(x.shape, np.complex64)
    xp.array[...] = x
    out["pinned_input_ms"] = wall(lambda: streamed_peaks(xp.array))
    out["pinned_h2d_GBps"] = x.nbytes / (wall(lambda: ctx.check(
        ctx.lib.spx_memcpy_h2d(ctx.h, d.ptr, xp.ptr, x.nbytes), "spx_memcpy_h2d")) * 1e-3) / 1e9
    xp.free()
    out["streamed_channels_per_s"] = B / (out["streamed_ms"] * 1e-3)
    spec = ctx.alloc(x.nbytes)
    host_spec = np.empty_like(x)

    def plain_c2c():
        d.upload(x)
        ctx.fft_c2c(d, spec, B, N, N, 0)
        ctx.sync()
        spec.download(np.complex64, x.shape)

    out["c2c_plain_ms"] = wall(plain_c2c)
    out["c2c_streamed_ms"] = wall(lambda: ctx.fft_c2c_host(x, N, out=host_spec))
    for b in (d, pk, ct, g, spec):
        b.free()
    return out


def extra_uneven(ctx, K=64):
    import tempfile

    from spectrum_amd import capi

    n = 4096
    L = capi.siggen_lengths(0x5EED0003, n)
    off, total, groups = capi.length_major_offsets(L)
    x = capi.siggen_uneven(capi.GEN_MULTI_TONE, 0x5EED0003, L, off, total)
    d = ctx.to_device(x)
    pk, ct = ctx.alloc(n * K * 32), ctx.alloc(n * 8)
    # the autotuner times every candidate of every length on this batch and
    # writes its cache; the measured plan reads it (the untuned plan beside it)
    cache = os.path.join(tempfile.mkdtemp(prefix="spx_tune_"), "uneven.json")
    t0 = time.perf_counter()
    ctx.batch_autotune(ctx.batch_desc(d, off, L, pk, ct, max_peaks=K), cache)
    tune_s = time.perf_counter() - t0
    tuned = json.load(open(cache))
    plan = ctx.batch_plan(ctx.batch_desc(d, off, L, pk, ct, max_peaks=K, tune_cache=cache))
    plan0 = ctx.batch_plan(ctx.batch_desc(d, off, L, pk, ct, max_peaks=K))
    bufs = {Lx: (ctx.alloc(len(ids) * K * 32), ctx.alloc(len(ids) * 8)) for Lx, (o, ids) in groups.items()}

    def seq():
        for Lx, (o, ids) in sorted(groups.items()):
            ctx.fused_peaks(d.ptr + 8 * o, len(ids), Lx, Lx, bufs[Lx][0], bufs[Lx][1], max_peaks=K)

    # the forms alternate over three rounds (the first timed round of a run
    # reads slow), medians reported
    tb, t0s, ts = [], [], []
    for _ in range(3):
        tb.append(_timer_ms(ctx, plan.execute))
        t0s.append(_timer_ms(ctx, plan0.execute))
        ts.append(_timer_ms(ctx, seq))
    t_batch, t_untuned, t_seq = float(np.median(tb)), float(np.median(t0s)), float(np.median(ts))
    info = plan.info()
    plan.close()
    plan0.close()
    alg = total * 8 + n * (K * 32 + 8)
    return {"workload": f"{n} channels, lengths 2^10..2^17 log-uniform (complex64, {total * 8 / 1e9:.2f} GB), "
                        f"Hann, mean+6std, K={K}", "batch_graph_ms": t_batch, "untuned_graph_ms": t_untuned,
            "per_length_launches_ms": t_seq, "speedup": t_seq / t_batch, "channels_per_s": n / (t_batch * 1e-3),
            "achieved_GBps": alg / (t_batch * 1e-3) / 1e9, "autotune_s": tune_s,
            "tuned": [{k: e[k] for k in ("length", "order", "tile", "groups", "us")} for e in tuned["lengths"]],
            "concurrent": bool(tuned.get("concurrent", 0)), "tile_sets": tuned.get("tile_sets", []),
            "plan": info}


def _rmse_vs_tones(recs, cnt, tf, ta, N):
    """RMS distance (bins) from each stored peak's refined frequency to the
    nearest true tone, over peaks within one bin of a tone."""
    errs = []
    for b in range(recs.shape[0]):
        f = recs["refined_frequency"][b, : cnt[b]].astype(np.float64) * N
        tb = np.mod(tf[b][ta[b] > 0] * N, N)
        if len(f) and len(tb):
            dd = (f[:, None] - tb[None, :] + N / 2) % N - N / 2
            e = dd[np.arange(len(f)), np.argmin(np.abs(dd), axis=1)]
            errs.append(e[np.abs(e) <= 1.0])
    e = np.concatenate(errs) if errs else np.zeros(0)
    return float(np.sqrt(np.mean(e * e))) if len(e) else None, int(len(e))


def extra_mixed_mlp(ctx, K=32, reps=30):
    """C4: 2048 x 16384 stress signal.  The FP16/BF16-MFMA form (learned and
    parabolic refinement) against the plain FP32 form on complex64 input;
    spectral error of the MFMA transform against FP64 (transform: FP64 FFT of
    the stored input; end to end: of the complex64 signal before storage);
    peak-index agreement over all channels (exact lists vs the FP32 form, and
    the tie rule vs the oracle on the FP64 spectrum of the stored input);
    refined-frequency RMSE against the generator's true tones."""
    from spectrum_amd import capi
    from spectrum_amd.refiner import DEFAULT_WEIGHTS, Refiner
    from oracle import oracle as orc

    B, N = 2048, 16384
    x, tf, ta = capi.siggen(capi.GEN_DYNAMIC, 0x5EED0004, B, N, max_tones=2, with_tones=True)
    fs = 1.0
    d64 = ctx.to_device(x)
    pk, ct = ctx.alloc(B * K * 32), ctx.alloc(B * 8)
    out = {"workload": f"{B} x {N}, strong tone + tone at -60 dB + noise at -80 dB, Hann, mean+6std, K={K}"}
    r = Refiner(ctx).load(DEFAULT_WEIGHTS)
    r.export()
    dts = (("fp16", capi.DTYPE_C32F16), ("bf16", capi.DTYPE_C32BF16))
    stored = {name: capi.to_half_storage(x, dt) for name, dt in dts}
    dev = {name: ctx.to_device(stored[name]) for name, _ in dts}

    def run_mixed(name, dt, learned):
        ctx.set_refiner(r.deployed if learned else None)
        ctx.fused_peaks_mixed(dev[name], B, N, pk, ct, dtype=dt, max_peaks=K, window=1,
                              refine=capi.REFINE_LEARNED if learned else capi.REFINE_PARABOLIC)

    # every form is timed in each of three rounds, in turn (the first timed
    # call of a run reads slow), and the medians are reported
    forms = [("fp32", lambda: ctx.fused_peaks(d64, B, N, N, pk, ct, max_peaks=K, window=1))]
    for name, dt in dts:
        forms.append((name + "_learned", lambda name=name, dt=dt: run_mixed(name, dt, True)))
        forms.append((name + "_parabolic", lambda name=name, dt=dt: run_mixed(name, dt, False)))
    times = {f: [] for f, _ in forms}
    for _ in range(3):
        for f, fn in forms:
            times[f].append(_timer_ms(ctx, fn, reps))
    ctx.set_refiner(None)
    med = {f: float(np.median(v)) for f, v in times.items()}
    t_fp32 = med["fp32"]
    out["fp32_c64_parabolic_ms"] = t_fp32
    ctx.fused_peaks(d64, B, N, N, pk, ct, max_peaks=K, window=1)
    ctx.sync()
    rec32 = pk.download(capi.MAXVALUE_DTYPE, (B, K))
    cnt32 = ct.download(np.uint32, (B, 2))
    out["fp32_rmse_bins"], _ = _rmse_vs_tones(rec32, cnt32[:, 0], tf, ta, N)
    S = 64  # channels whose spectra are compared against FP64
    for name, dt in dts:
        u, d16 = stored[name], dev[name]
        t_l, t_p = med[name + "_learned"], med[name + "_parabolic"]
        run_mixed(name, dt, True)
        ctx.sync()
        recl = pk.download(capi.MAXVALUE_DTYPE, (B, K))
        cntl = ct.download(np.uint32, (B, 2))
        run_mixed(name, dt, False)
        ctx.set_refiner(None)
        ctx.sync()
        recp = pk.download(capi.MAXVALUE_DTYPE, (B, K))
        cntp = ct.download(np.uint32, (B, 2))
        # the FP32 form on the same stored input: agreement judged on one storage
        ctx.fused_peaks(d16, B, N, N, pk, ct, dtype=dt, max_peaks=K, window=1)
        ctx.sync()
        recf = pk.download(capi.MAXVALUE_DTYPE, (B, K))
        cntf = ct.download(np.uint32, (B, 2))
        same_fp32 = sum(list(recp[b, : cntp[b, 0]]["index"]) == list(recf[b, : cntf[b, 0]]["index"])
                        for b in range(B))
        xq = capi.from_half_storage(u, dt)
        Xq = orc.spectrum64(xq, N, 1)
        ok_oracle = 0
        for b in range(B):
            recs, found, g, stats, mags = orc.fused_peaks(Xq[b], fs, nsigma=6.0, kcap=K)
            ok, _ = orc.compare_peak_lists(recp[b, : cntp[b, 0]]["index"], recs["index"], mags, stats[2],
                                           2e-3 * mags.max(), K)
            ok_oracle += int(ok)
        sp = ctx.alloc(S * N * 8)
        ctx.fft_mixed(d16, sp, S, N, window=1, dtype=dt)
        ctx.sync()
        X = sp.download(np.complex64, (S, N))
        e_t = np.linalg.norm(X - Xq[:S], axis=1) / np.linalg.norm(Xq[:S], axis=1)
        ref = orc.spectrum64(x[:S], N, 1)
        e_e = np.linalg.norm(X - ref, axis=1) / np.linalg.norm(ref, axis=1)
        e_s = np.linalg.norm(xq[:S] - x[:S], axis=1) / np.linalg.norm(x[:S], axis=1)
        rl, nl = _rmse_vs_tones(recl, cntl[:, 0], tf, ta, N)
        rp, _ = _rmse_vs_tones(recp, cntp[:, 0], tf, ta, N)
        out[name] = {"mfma_learned_ms": t_l, "mfma_parabolic_ms": t_p, "speedup_vs_fp32": t_fp32 / t_l,
                     "speedup_vs_fp32_parabolic": t_fp32 / t_p,
                     "transform_rel_l2_max": float(e_t.max()), "end_to_end_rel_l2_max": float(e_e.max()),
                     "storage_rel_l2_max": float(e_s.max()),
                     "peak_lists_equal_fp32_form": f"{same_fp32}/{B}",
                     "peak_lists_tie_rule_vs_oracle": f"{ok_oracle}/{B}",
                     "learned_rmse_bins": rl, "parabolic_rmse_bins": rp, "rmse_peaks": nl}
        for buf in (d16, sp):
            buf.free()
    out["speedup_vs_fp32"] = out["fp16"]["speedup_vs_fp32"]
    out["error_reference"] = ("transform: FP64 FFT of the stored input over " + str(S) + " channels; end to end: "
                              "FP64 FFT of the complex64 signal before storage")
    mf = os.path.join(ROOT, "profiles", "mixed16k_counters.json")
    if os.path.exists(mf):
        out["mfma_util"] = json.load(open(mf)).get("mfma_util")
    return out


def extra_refiner(ctx, K=64):
    from spectrum_amd import capi
    from spectrum_amd.refiner import DEFAULT_WEIGHTS, Refiner

    B, N = 4096, 8192
    x, tf, ta = capi.siggen(capi.GEN_FRACTIONAL, 0x5EED0005, B, N, max_tones=16, with_tones=True)
    d = ctx.to_device(x)
    pk, ct = ctx.alloc(B * K * 32), ctx.alloc(B * 8)
    r = Refiner(ctx).load(DEFAULT_WEIGHTS)
    r.export()
    out = {"workload": f"{B} x {N} complex64, 1-16 fractional-bin tones, Hann, mean+6std, K={K}"}
    for name, refine in (("learned", capi.REFINE_LEARNED), ("parabolic", capi.REFINE_PARABOLIC)):
        ctx.set_refiner(r.deployed if refine == capi.REFINE_LEARNED else None)
        out[f"{name}_ms"] = _timer_ms(ctx, lambda: ctx.fused_peaks(d, B, N, N, pk, ct, max_peaks=K, refine=refine))
        recs = pk.download(capi.MAXVALUE_DTYPE, (B, K))
        cnt = ct.download(np.uint32, (B, 2))[:, 0]
        errs = []
        for b in range(B):
            f = recs["refined_frequency"][b, : cnt[b]].astype(np.float64) * N
            tb = np.mod(tf[b][ta[b] > 0] * N, N)
            if len(f) and len(tb):
                dd = (f[:, None] - tb[None, :] + N / 2) % N - N / 2
                e = dd[np.arange(len(f)), np.argmin(np.abs(dd), axis=1)]
                errs.append(e[np.abs(e) <= 1.0])
        e = np.concatenate(errs)
        out[f"{name}_rmse_bins"] = float(np.sqrt(np.mean(e * e)))
    ctx.set_refiner(None)
    return out


def extra_refiner_split(ctx, K=64, reps=20):
    """Learned refinement on the split path at the headline shape: 1024 x 65536
    complex64 fractional-bin tones (1-16 per channel), Hann, mean+6std, K=64;
    time per call and frequency RMSE against the true tones, learned vs
    parabolic."""
    from spectrum_amd import capi
    from spectrum_amd.refiner import DEFAULT_WEIGHTS, Refiner

    B, N = 1024, 65536
    x, tf, ta = capi.siggen(capi.GEN_FRACTIONAL, 0x5EED0006, B, N, max_tones=16, with_tones=True)
    d = ctx.to_device(x)
    pk, ct = ctx.alloc(B * K * 32), ctx.alloc(B * 8)
    r = Refiner(ctx).load(DEFAULT_WEIGHTS)
    r.export()
    out = {"workload": f"{B} x {N} complex64, 1-16 fractional-bin tones, Hann, mean+6std, K={K} (split path)"}
    for name, refine in (("learned", capi.REFINE_LEARNED), ("parabolic", capi.REFINE_PARABOLIC)):
        ctx.set_refiner(r.deployed if refine == capi.REFINE_LEARNED else None)
        out[f"{name}_ms"] = _timer_ms(ctx, lambda: ctx.fused_peaks(d, B, N, N, pk, ct, max_peaks=K, refine=refine),
                                      reps)
        recs = pk.download(capi.MAXVALUE_DTYPE, (B, K))
        cnt = ct.download(np.uint32, (B, 2))[:, 0]
        out[f"{name}_rmse_bins"], out["rmse_peaks"] = _rmse_vs_tones(recs, cnt, tf, ta, N)
    ctx.set_refiner(None)
    return out


def extra_small_batch(ctx, K=32, reps=100):
    """64 x 4096 complex64, three tones at 10 dB (the CPU-runnable shape):
    one fused launch per call, latency-bound; beside it the reference CPU path
    (pocketfft float via the reference CpuFFT, one thread, the same peak rules)
    on all 64 channels, 1 warm-up run, median of 5."""
    from spectrum_amd import capi
    from oracle import oracle as orc

    B, N = 64, 4096
    x = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0001, B, N)
    d = ctx.to_device(x)
    pk, ct, gm = ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32)
    ms = _timer_ms(ctx, lambda: ctx.fused_peaks(d, B, N, N, pk, ct, max_peaks=K, gmax=gm), reps)
    out = {"workload": f"{B} x {N} complex64, 3 tones at 10 dB, Hann, mean+6std, K={K}, global max",
           "ms": ms, "channels_per_s": B / (ms * 1e-3),
           "achieved_GBps": B * (8 * N + K * 32 + 8 + 32) / (ms * 1e-3) / 1e9}
    w = orc.window(1, N).astype(np.float32)
    kind = "reference" if orc.have_reference() else "port"

    def cpu():
        for b in range(B):
            xw = (x[b] * w).astype(np.complex64)[None]
            X = orc.ref_cpufft(xw)[0] if kind == "reference" else orc.spectrum64(xw, N, 0)[0].astype(np.complex64)
            orc.fused_peaks(X, 1.0, nsigma=6.0, kcap=K)

    med, ts = _median_runs(cpu, 5)
    out["cpu_baseline"] = {"ms": med * 1e3, "channels_per_s": B / med, "cores": 1, "kind": kind}
    out["gpu_vs_cpu"] = med * 1e3 / ms
    return out


def extra_large(ctx, K=64, reps=20):
    """Lengths past the split path: 64 x 2^20 complex64 (512 MB), Hann,
    mean+6std, K=64, and the complex spectrum of the same batch.  Three
    Stockham passes through HBM (a radix-16 register pass, two 256-point
    LDS-tile passes), the threshold statistics inside the last pass, then the
    candidate search on the stored spectrum; algorithmic bytes are the input
    (plus the spectrum for the dense output)."""
    from spectrum_amd import capi

    B, N = 64, 1 << 20
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0314, B, N)
    d = ctx.to_device(x)
    out = ctx.alloc(B * N * 8)
    pk, ct = ctx.alloc(B * K * 32), ctx.alloc(B * 8)
    t_p = _timer_ms(ctx, lambda: ctx.fused_peaks(d, B, N, N, pk, ct, max_peaks=K), reps, 5)
    t_c = _timer_ms(ctx, lambda: ctx.fft_c2c(d, out, B, N, N, 1), reps, 5)
    alg_p = B * (8 * N + K * 32 + 8)
    alg_c = B * 16 * N
    return {"workload": f"{B} x {N} complex64, Hann, mean+6std, K={K}; and its complex spectrum",
            "peaks_ms": t_p, "peaks_channels_per_s": B / (t_p * 1e-3),
            "peaks_achieved_GBps": alg_p / (t_p * 1e-3) / 1e9, "peaks_frac": alg_p / (t_p * 1e-3) / 8e12,
            "c2c_ms": t_c, "c2c_achieved_GBps": alg_c / (t_c * 1e-3) / 1e9, "c2c_frac": alg_c / (t_c * 1e-3) / 8e12}


def extra_top_peaks(ctx, reps=20):
    """The reference ONE_PEAK / TWO_PEAKS search of SpectrumProcessorROCm at
    1024 x 65536 complex64 (search range nfft/4): inside the transform
    (spx_fused_peaks_top: three launches, no spectrum in HBM) against the
    reference's chain, transform to a stored spectrum then the search
    (spx_fft_c2c + spx_peaks_top)."""
    from spectrum_amd import capi

    B, N = 1024, 65536
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0007, B, N)
    d = ctx.to_device(x)
    spec = ctx.alloc(B * N * 8)
    out = {"workload": f"{B} x {N} complex64, search range {N // 4}, fs 1"}
    for name, two in (("one_peak", False), ("two_peaks", True)):
        R = 8 if two else 4
        o1, o2 = ctx.alloc(B * R * 32), ctx.alloc(B * R * 32)
        t_f = _timer_ms(ctx, lambda: ctx.fused_peaks_top(d, o1, B, N, N, N // 4, 1.0, two_peaks=two), reps)

        def chain():
            ctx.fft_c2c(d, spec, B, N, N)
            ctx.peaks_top(spec, o2, B, N, N // 4, 1.0, two_peaks=two)

        t_c = _timer_ms(ctx, chain, reps)
        a = o1.download(capi.MAXVALUE_DTYPE, (B, R))
        b = o2.download(capi.MAXVALUE_DTYPE, (B, R))
        out[name] = {"fused_ms": t_f, "fft_then_search_ms": t_c, "speedup": t_c / t_f,
                     "records_equal": bool(np.array_equal(a["index"], b["index"]))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: the first few dozen calls after start-up run ~10 % slower
    # (clocks ramping up); 50 warm-up calls reach the steady state
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--nfft", type=int, default=65536)
    ap.add_argument("--kpeaks", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--extra", choices=["host_stream", "uneven", "mixed_mlp", "refiner", "refiner_split", "top_peaks",
                                        "small_batch", "large"],
                    help="run only this extra workload (no headline, no CPU baseline) and print its JSON; "
                         "for profiling one workload")
    ap.add_argument("--reps", type=int, default=30, help="timed repetitions of each --extra form")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after timing, copy the input once device-to-device (a plain copy of known size, "
                         "used to scale rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported (channels shard "
                                   "across GPUs with no exchange; see DESIGN.md)"}))
        sys.exit(2)

    from spectrum_amd import capi

    B, N, K = args.batch, args.nfft, args.kpeaks
    ctx = capi.Context(0)
    if args.extra:
        if args.extra == "host_stream":
            res = extra_host(ctx, capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0002, B, N), K)
        elif args.extra == "mixed_mlp":
            res = extra_mixed_mlp(ctx, reps=args.reps)
        else:
            res = {"uneven": extra_uneven, "refiner": extra_refiner, "refiner_split": extra_refiner_split,
                   "top_peaks": extra_top_peaks, "small_batch": extra_small_batch, "large": extra_large}[args.extra](ctx)
        print(json.dumps({args.extra: res}), flush=True)
        return
    info = ctx.info()
    print(f"[bench] device {info}", file=sys.stderr, flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0002, B, N)
    d_in = ctx.to_device(x)
    d_pk = ctx.alloc(B * K * 32)
    d_ct = ctx.alloc(B * 8)
    d_g = ctx.alloc(B * 32)

    def step():
        ctx.fused_peaks(d_in, B, N, N, d_pk, d_ct, window=1, fs=1.0, nsigma=6.0, max_peaks=K, gmax=d_g)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    timer = HipTimer(ctx.stream())
    per = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        timer.start()
        step()
        per.append(timer.stop_ms())
    ctx.sync()
    wall = time.perf_counter() - t0
    ms = float(np.mean(per))
    print(f"[bench] per-step ms: mean {ms:.4f} median {np.median(per):.4f} min {np.min(per):.4f}",
          file=sys.stderr, flush=True)

    alg_bytes = B * (8 * N + K * 32 + 8 + 32)
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "fourstep_n65536_b1024_traffic.json")
    if (B, N, K) == (1024, 65536, 64) and os.path.exists(tf):
        traffic = json.load(open(tf)).get("hbm_bytes_per_step")
    if args.calibrate_copy:
        hip = C.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        d_cp = ctx.alloc(x.nbytes)
        assert hip.hipMemcpy(d_cp.ptr, d_in.ptr, x.nbytes, 3) == 0  # hipMemcpyDeviceToDevice
        ctx.sync()
        d_cp.free()
    ct = d_ct.download(np.uint32, (B, 2))
    out = {
        "metric": METRIC,
        "value": B / (ms * 1e-3),
        "unit": "channels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"fused window->FFT->|X|->peaks, {B}x{N} complex64, Hann, mean+6std, K={K}, "
                               "parabolic, global max", "batch": B, "nfft": N, "max_peaks": K,
                   "peaks_found_mean": float(ct[:, 1].mean())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernels": "k_fs_cols + k_fs_rows + k_fs_finish (one call)"},
        "wall_s_timed": wall,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, N, K)
        try:
            mt = cpu_baseline_mt(x, N, K, min(16, os.cpu_count() or 1))
            if mt:
                out["cpu_baseline"]["multicore"] = mt
        except Exception as e:  # the single-thread baseline stands on its own
            out["cpu_baseline"]["multicore"] = {"error": repr(e)}
    if not args.no_extras:
        print("[bench] host_stream ...", file=sys.stderr, flush=True)
        try:
            out["host_stream"] = extra_host(ctx, x, K)
        except Exception as e:
            out["host_stream"] = {"error": repr(e)}
        del d_in, x
        for name, fn in (("small_batch", extra_small_batch), ("uneven", extra_uneven), ("mixed_mlp", extra_mixed_mlp),
                         ("refiner", extra_refiner), ("refiner_split", extra_refiner_split),
                         ("top_peaks", extra_top_peaks), ("large", extra_large)):
            print(f"[bench] {name} ...", file=sys.stderr, flush=True)
            try:
                out[name] = fn(ctx)
            except Exception as e:
                out[name] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
