"""CPU checks that pin the refiner oracle (oracle/mlp.py) before it judges the GPU kernels."""
import numpy as np

from oracle import mlp
from oracle import oracle as orc


def _params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp.NPARAM)
    p[mlp.OFF_W1:mlp.OFF_B1] = rng.uniform(-1, 1, mlp.H * mlp.IN) * np.sqrt(6 / mlp.IN)
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.normal(0, 0.1, mlp.H)
    p[mlp.OFF_W2:mlp.OFF_B2] = rng.uniform(-1, 1, mlp.H * mlp.H) * np.sqrt(6 / mlp.H)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.normal(0, 0.1, mlp.H)
    p[mlp.OFF_W3:mlp.OFF_B3] = rng.uniform(-1, 1, mlp.H) * np.sqrt(6 / mlp.H)
    p[mlp.OFF_B3] = 0.05
    return p


def test_param_count():
    assert mlp.NPARAM == 4929


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(1)
    X = rng.normal(0, 1, (37, mlp.IN))
    y = rng.normal(0, 0.3, 37)
    g = mlp.grad(p, X, y)
    idx = np.concatenate([rng.choice(mlp.NPARAM, 60, replace=False), [mlp.OFF_B3, mlp.OFF_W1, mlp.OFF_B2]])
    fd = mlp.grad_fd(p, X, y, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_planted_mistake_is_caught():
    """A dropped ReLU mask in the backward pass must fail the finite-difference check."""
    rng = np.random.default_rng(2)
    p = _params(2)
    X = rng.normal(0, 1, (20, mlp.IN))
    y = rng.normal(0, 0.3, 20)
    W1, b1, W2, b2, w3, b3 = mlp.unpack(p)
    out, (Xf, a1, h1, a2, h2) = mlp.forward(p, X)
    dout = 2.0 * (out - y) / len(y)
    d2_wrong = np.outer(dout, w3)  # missing (a2 > 0)
    gw = (d2_wrong.T @ h1).ravel()
    idx = np.arange(mlp.OFF_W2, mlp.OFF_W2 + 200)
    fd = mlp.grad_fd(p, X, y, idx)
    assert not np.allclose(gw[:200], fd, rtol=1e-3, atol=1e-6)


def test_features_closed_form():
    """Integer-bin tone, rectangular window: X is a single spike, so z = (0, 0, 1, 0, 0)."""
    n, k0 = 256, 40
    x = np.exp(2j * np.pi * k0 * np.arange(n) / n) * (0.3 - 0.4j)
    X = orc.fft64(x)
    f = mlp.features(X, k0)
    assert np.allclose(f, [0, 0, 0, 0, 1, 0, 0, 0, 0, 0], atol=1e-12)


def test_label_recovers_true_frequency():
    """Hann-windowed fractional tone: the 16x padded label is within 3e-3 bin of the truth."""
    n = 8192
    w = orc.window(1, n)
    for k0, u in ((1000, 0.1), (2345, 0.37), (7000, 0.5), (100, 0.93)):
        f = (k0 + u) / n
        x = np.exp(2j * np.pi * f * np.arange(n)) * w
        xp = np.zeros(16 * n, np.complex128)
        xp[:n] = x
        m16 = np.abs(orc.fft64(xp))
        X = orc.fft64(x)
        k = int(np.argmax(np.abs(X)))
        lab = mlp.label(m16, k)
        assert lab is not None
        assert abs((k + lab) - (k0 + u)) < 3e-3, (k0, u, k, lab)
