"""Headline benchmark: fused window -> FFT -> |X| -> peak search on one MI355X.

Workload (the metric's configuration): 1024 channels x 65536-point complex64,
1-8 tones per channel at 0-20 dB SNR, Hann window, threshold mean + 6 std of
|X|, first 64 local maxima in ascending bin order, parabolic refinement,
global-maximum record.  One step = one launch over the whole batch, input
already resident in HBM.

Prints one JSON line (see README of the driver contract): value = channels/s,
roofline = algorithmic bytes / average kernel time vs 8 TB/s, cpu_baseline =
the reference's CPU path (its pocketfft CpuFFT, compiled from the reference
sources into oracle/_ref) plus the oracle's C peak rules, timed on a bounded
sample on this host.
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


def cpu_baseline(x, nfft, K, budget_s=10.0):
    """Reference CPU FFT (pocketfft float, one thread, plan per call) + oracle peak rules."""
    from oracle import oracle as orc

    w = orc.window(1, nfft).astype(np.float32)
    kind = "reference" if orc.have_reference() else "port"
    done = 0
    t0 = time.perf_counter()
    for b in range(x.shape[0]):
        xw = (x[b] * w).astype(np.complex64)[None]
        X = orc.ref_cpufft(xw)[0] if kind == "reference" else orc.spectrum64(xw, nfft, 0)[0].astype(np.complex64)
        orc.fused_peaks(X, 1.0, nsigma=6.0, kcap=K)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {
        "value": done / dt,
        "unit": "channels/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{done} of {x.shape[0]} channels x {nfft} c64 (Hann, pocketfft c2c<float> via the reference "
                  f"CpuFFT::ForwardC2C, then |X|/mean+6std/local-max/K={K}/parabolic), single thread, "
                  f"host {os.cpu_count()} logical cpus",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--nfft", type=int, default=65536)
    ap.add_argument("--kpeaks", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
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
    tf = os.path.join(ROOT, "profiles", "fused_large_n65536_b1024_traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "wall_s_timed": wall,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, N, K)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
