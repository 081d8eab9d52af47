"""Learned sub-bin refiner: data generation, GPU training and deployment.

The network (10 -> 64 -> 64 -> 1, ReLU) reads the 5-bin complex
neighbourhood of a detected peak, X[k+j] / X[k] for j = -2..2, and predicts
the tone's offset from bin k.  Everything numeric runs on the GPU through the
C ABI: signals come from the library's generator (1-16 tones at fractional
bins, Hann window), peaks from spx_fused_peaks, labels from the 16x
zero-padded transform of the same windowed signal (spx_fft_magnitude with
nfft = 16 N) refined by a parabola on the fine grid, and training uses the
library's forward/backward and Adam kernels.  The deployed block (FP16
weights) is what spx_fused_peaks uses with refine = REFINE_LEARNED.
"""
from __future__ import annotations

import hashlib
import os
from dataclasses import dataclass

import numpy as np

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "refiner_hann_8192.npz")

# Weights file: an .npz (no pickled objects) with the FP32 master parameters
# `params` in the order the library's kernels read them, the architecture tag
# `format`, the file-format `version` and the SHA-256 of the parameters'
# little-endian bytes (`sha256`), plus free-form training metadata.
WEIGHTS_FORMAT = "spx-refiner mlp 10-64-64-1 relu fp32"
WEIGHTS_VERSION = 1
WEIGHTS_NPARAM = 10 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def params_digest(p: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(p, dtype="<f4").tobytes()).hexdigest()


def read_weights(path: str) -> np.ndarray:
    """Parameters of a weights file, checked: architecture tag and version
    (when the file has them), parameter count and dtype, finiteness, and the
    checksum.  Raises ValueError on any mismatch."""
    with np.load(path, allow_pickle=False) as z:
        if "params" not in z.files:
            raise ValueError(f"{path}: no 'params' array")
        p = z["params"]
        if "format" in z.files and str(z["format"]) != WEIGHTS_FORMAT:
            raise ValueError(f"{path}: format {str(z['format'])!r}, expected {WEIGHTS_FORMAT!r}")
        if "version" in z.files and int(z["version"]) > WEIGHTS_VERSION:
            raise ValueError(f"{path}: file version {int(z['version'])} is newer than {WEIGHTS_VERSION}")
        digest = str(z["sha256"]) if "sha256" in z.files else None
    if p.dtype != np.float32 or p.shape != (WEIGHTS_NPARAM,):
        raise ValueError(f"{path}: params {p.dtype}{p.shape}, expected float32({WEIGHTS_NPARAM},)")
    if not np.all(np.isfinite(p)):
        raise ValueError(f"{path}: non-finite parameters")
    if digest is not None and digest != params_digest(p):
        raise ValueError(f"{path}: checksum mismatch (file damaged or edited)")
    return p


def write_weights(path: str, params: np.ndarray, **meta) -> None:
    p = np.ascontiguousarray(params, dtype=np.float32)
    if p.shape != (WEIGHTS_NPARAM,):
        raise ValueError(f"params shape {p.shape}, expected ({WEIGHTS_NPARAM},)")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=p, format=np.asarray(WEIGHTS_FORMAT), version=np.asarray(WEIGHTS_VERSION),
             sha256=np.asarray(params_digest(p)), **{k: np.asarray(v) for k, v in meta.items()})


@dataclass
class PeakSet:
    feat: np.ndarray        # [n, 10] float32
    label: np.ndarray       # [n] offset from the 16x zero-padded FFT (bins)
    parab: np.ndarray       # [n] parabolic offset of the same peak
    truth: np.ndarray       # [n] offset from the generator's nearest tone (bins), nan if none within 1 bin
    channel: np.ndarray     # [n] generator channel index
    slot: np.ndarray        # [n] record slot
    bin: np.ndarray         # [n] peak bin k

    def __len__(self):
        return len(self.label)


def make_peaks(ctx: "capi.Context", n_chan: int, seed: int, N: int = 8192, K: int = 64, chunk: int = 512,
               kind: int = capi.GEN_FRACTIONAL, chan_offset: int = 0, max_tones: int = 16) -> PeakSet:
    """Detected peaks of n_chan generated channels with features and labels."""
    feats, labels, parabs, truths, chans, slots, bins = [], [], [], [], [], [], []
    d_sig = ctx.alloc(chunk * N * 8)
    d_spec = ctx.alloc(chunk * N * 8)
    d_m16 = ctx.alloc(chunk * 16 * N * 4)
    d_pk, d_ct = ctx.alloc(chunk * K * 32), ctx.alloc(chunk * 8)
    d_f, d_l, d_p, d_v = ctx.alloc(chunk * K * 40), ctx.alloc(chunk * K * 4), ctx.alloc(chunk * K * 4), \
        ctx.alloc(chunk * K * 4)
    for c0 in range(0, n_chan, chunk):
        B = min(chunk, n_chan - c0)
        x, tf, ta = capi.siggen(kind, seed, B, N, chan_offset=chan_offset + c0, max_tones=max_tones,
                                with_tones=True)
        d_sig.upload(x)
        ctx.fused_peaks(d_sig, B, N, N, d_pk, d_ct, window=1, fs=1.0, nsigma=6.0, max_peaks=K,
                        refine=capi.REFINE_PARABOLIC)
        ctx.fft_c2c(d_sig, d_spec, B, N, N, 1)
        ctx.fft_magnitude(d_sig, d_m16, B, N, 16 * N, 1)
        ctx.mlp_dataset(d_spec, d_m16, d_pk, d_ct, B, N, K, d_f, d_l, d_p, d_v)
        ctx.sync()
        v = d_v.download(np.uint32, (B, K)).astype(bool)
        f = d_f.download(np.float32, (B, K, 10))
        lab = d_l.download(np.float32, (B, K))
        par = d_p.download(np.float32, (B, K))
        pk = d_pk.download(capi.MAXVALUE_DTYPE, (B, K))
        b_idx, s_idx = np.nonzero(v)
        feats.append(f[b_idx, s_idx])
        labels.append(lab[b_idx, s_idx])
        parabs.append(par[b_idx, s_idx])
        k = pk["index"][b_idx, s_idx].astype(np.float64)
        # generator truth: nearest tone (cycles/sample -> bins), within one bin
        fb = np.mod(tf[b_idx] * N, N)
        has = ta[b_idx] > 0
        dist = np.where(has, (fb - k[:, None] + N / 2) % N - N / 2, np.inf)
        j = np.argmin(np.abs(dist), axis=1)
        dt = dist[np.arange(len(j)), j]
        truths.append(np.where(np.abs(dt) <= 1.0, dt, np.nan).astype(np.float32))
        chans.append(b_idx + c0 + chan_offset)
        slots.append(s_idx)
        bins.append(pk["index"][b_idx, s_idx])
    for b in (d_sig, d_spec, d_m16, d_pk, d_ct, d_f, d_l, d_p, d_v):
        b.free()
    cat = np.concatenate
    return PeakSet(cat(feats), cat(labels), cat(parabs), cat(truths), cat(chans), cat(slots), cat(bins))


class Refiner:
    """FP32 master weights + Adam state on the device, deployed FP16 block for inference."""

    def __init__(self, ctx: "capi.Context", seed: int = 1):
        self.ctx = ctx
        self.P = ctx.lib.spx_mlp_param_count()
        self.params = ctx.alloc(self.P * 4)
        self.m = ctx.alloc(self.P * 4)
        self.v = ctx.alloc(self.P * 4)
        self.grad = ctx.alloc((self.P + 1) * 4)
        self.deployed = ctx.alloc(ctx.lib.spx_mlp_deployed_size())
        self.steps = 0
        ctx.mlp_init(self.params, seed)
        self._zero_adam()

    def _zero_adam(self):
        z = np.zeros(self.P, np.float32)
        self.m.upload(z)
        self.v.upload(z)
        self.steps = 0

    def set_params(self, p: np.ndarray):
        self.params.upload(np.ascontiguousarray(p, dtype=np.float32))
        self._zero_adam()

    def get_params(self) -> np.ndarray:
        return self.params.download(np.float32, (self.P,))

    def train(self, X: np.ndarray, Y: np.ndarray, *, steps: int, batch: int = 1024, lr: float = 1e-3,
              seed: int = 7) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        dX, dY = self.ctx.to_device(X), self.ctx.to_device(Y)
        losses = self.ctx.mlp_train(self.params, self.m, self.v, self.grad, dX, dY, len(Y), min(batch, len(Y)),
                                    steps, self.steps, lr, seed)
        self.steps += steps
        dX.free()
        dY.free()
        return losses

    def export(self):
        self.ctx.mlp_export(self.params, self.deployed)
        self.ctx.sync()
        return self.deployed

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        dX = self.ctx.to_device(X)
        out = self.ctx.alloc(max(len(X), 1) * 4)
        self.ctx.mlp_predict(self.deployed, dX, len(X), out)
        self.ctx.sync()
        r = out.download(np.float32, (len(X),))
        dX.free()
        out.free()
        return r

    def save(self, path: str, **meta):
        write_weights(path, self.get_params(), **meta)

    def load(self, path: str):
        self.set_params(read_weights(path))
        return self


def rmse(a, b) -> float:
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    d = d[np.isfinite(d)]
    return float(np.sqrt(np.mean(d * d))) if len(d) else float("nan")
