"""Pins the FP64 MLP restatement (oracle/mlp_oracle.py) on the CPU: BF16
rounding on known values, the layer wiring on a closed form, and the
analytic gradient on central finite differences."""
import numpy as np

from oracle import mlp_oracle as O
from libatssim_amd.warm_start import NUM_PARAMS, OFF_B4, OFF_W1, OFF_W2, OFF_W3, OFF_W4


def test_bf16_round_to_nearest_even():
    e7, e8 = 2.0 ** -7, 2.0 ** -8
    vals = np.array([1.0, 1 + e8, 1 + 3 * e8, 1 + e8 + 2.0 ** -20, -(1 + 3 * e8), 3.0, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 * e7, 1 + e7, -(1 + 2 * e7), 3.0, 0.0])
    assert np.array_equal(O.bf16(vals), want)


def test_wiring_closed_form():
    # one path x0 -> h1[0] -> h2[0] -> h3[0] -> y[2]: y2 = silu(silu(silu(x0)))
    p = np.zeros(NUM_PARAMS)
    p[OFF_W1 + 0 * 128 + 0] = 1.0
    p[OFF_W2 + 0 * 128 + 0] = 1.0
    p[OFF_W3 + 0 * 128 + 0] = 1.0
    p[OFF_W4 + 0 * 16 + 2] = 1.0
    p[OFF_B4 + 1] = 0.25
    x = np.zeros((3, 48))
    x[:, 0] = [-1.5, 0.3, 2.0]
    x[:, 5] = 7.0          # unused input
    y, _ = O.forward(p, x, bf16_round=False)
    s = lambda z: z / (1 + np.exp(-z))
    # the hidden units with zero weights still carry silu(0) = 0, so the path is exact
    assert np.allclose(y[:, 2], s(s(s(x[:, 0]))), rtol=1e-14, atol=0)
    assert np.allclose(y[:, 1], 0.25, rtol=0, atol=0)
    assert np.allclose(y[:, 0], 0.0, rtol=0, atol=0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = rng.standard_normal(NUM_PARAMS) * 0.1
    X = rng.standard_normal((9, 48))
    T = rng.standard_normal((9, 3))
    g, _ = O.gradient(p, X, T, bf16_round=False)
    idx = np.concatenate([rng.choice(NUM_PARAMS, 150, replace=False), np.arange(OFF_B4, OFF_B4 + 3),
                          [OFF_W1 + 5 * 128 + 7, OFF_W4 + 100 * 16 + 1]])
    h = 1e-6
    for i in idx:
        e = np.zeros(NUM_PARAMS)
        e[i] = h
        fd = (O.loss(p + e, X, T, False) - O.loss(p - e, X, T, False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
    # padded rows of W1 (inputs 48..63) and padded outputs 3..15 have zero gradient
    assert np.all(g[OFF_W1 + 48 * 128:OFF_W1 + 64 * 128] == 0)
    W4g = g[OFF_W4:OFF_B4].reshape(128, 16)
    assert np.all(W4g[:, 3:] == 0)


def test_bf16_gradient_is_close_to_exact():
    rng = np.random.default_rng(8)
    p = rng.standard_normal(NUM_PARAMS) * 0.1
    X = rng.standard_normal((64, 48))
    T = rng.standard_normal((64, 3))
    g0, _ = O.gradient(p, X, T, bf16_round=False)
    g1, _ = O.gradient(p, X, T, bf16_round=True)
    cos = g0 @ g1 / np.linalg.norm(g0) / np.linalg.norm(g1)
    assert cos > 0.999
