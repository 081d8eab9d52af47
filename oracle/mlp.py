"""FP64 CPU restatement of the learned sub-bin refiner -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's cpu_baseline leg may import this
module, and only as the checker.  It restates, in numpy float64:
  * the network 10 -> 64 -> 64 -> 1 (ReLU) and its mean-squared-error loss,
  * the analytic gradient (backpropagation), itself pinned by central finite
    differences (grad_fd),
  * the peak features z_j = X[k+j] / X[k], j = -2..2,
  * the label: position of the maximum of the 16x zero-padded |X| within
    +-11 fine bins of 16k, refined by a parabola, in coarse bins.
There is no reference implementation of this part of the path (the
reference refines peaks parabolically only,
include/spectrum/kernels/fft_kernel_sources_rocm.hpp:228-239); its parity
anchors are the finite-difference check and the generator's true tone
frequencies.
"""
from __future__ import annotations

import numpy as np

IN, H = 10, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    out = h2 @ w3 + b3
    return out, (X, a1, h1, a2, h2)


def loss(p, X, y):
    out, _ = forward(p, X)
    return float(np.mean((out - np.asarray(y, dtype=np.float64)) ** 2))


def grad(p, X, y):
    """Analytic gradient of the mean squared error (backpropagation)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    out, (X, a1, h1, a2, h2) = forward(p, X)
    n = X.shape[0]
    dout = 2.0 * (out - np.asarray(y, dtype=np.float64)) / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = h2.T @ dout
    g[OFF_B3] = dout.sum()
    d2 = np.outer(dout, w3) * (a2 > 0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(axis=0)
    d1 = (d2 @ W2) * (a1 > 0)
    g[OFF_W1:OFF_B1] = (d1.T @ X).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(axis=0)
    return g


def grad_fd(p, X, y, idx, h=1e-6):
    """Central finite differences of loss() at the parameter indices idx."""
    p = np.asarray(p, dtype=np.float64)
    out = []
    for i in idx:
        q1, q2 = p.copy(), p.copy()
        q1[i] += h
        q2[i] -= h
        out.append((loss(q1, X, y) - loss(q2, X, y)) / (2 * h))
    return np.array(out)


def features(X, k):
    """z_j = X[k+j] / X[k], j = -2..2 (circular), as 10 reals (re, im per j)."""
    X = np.asarray(X, dtype=np.complex128)
    n = X.shape[-1]
    c = X[k]
    z = np.array([X[(k + j) % n] / c for j in range(-2, 3)])
    return np.stack([z.real, z.imag], axis=1).ravel()


def label(mag16, k):
    """Offset (coarse bins) of the padded maximum near 16k, parabola-refined; None on the search edge."""
    m = np.asarray(mag16, dtype=np.float64)
    nf = m.shape[-1]
    o = np.arange(-11, 12)
    vals = m[(16 * k + o) % nf]
    b = int(np.argmax(vals))
    if b == 0 or b == len(o) - 1:
        return None
    yl, yc, yr = vals[b - 1], vals[b], vals[b + 1]
    den = yl - 2 * yc + yr
    d = 0.5 * (yl - yr) / den if den != 0 else 0.0
    return (o[b] + d) / 16.0
