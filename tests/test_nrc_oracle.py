"""The neural-radiance-cache oracle: analytic gradients against central finite differences (float64), the
bf16 emulation, and the Adam step against its formula."""
import numpy as np

from oracle import nrc_oracle as N


def test_gradients_match_finite_differences():
    assert N.check_gradients(0, B=8) < 1e-6
    assert N.check_gradients(3, B=16) < 1e-6


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 0.0], np.float32)
    r = N.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # 1 + 2^-8 is half an ulp of bf16: a tie, rounds to even (1.0)
    assert r[2] == 1.0                       # below half an ulp rounds down
    assert r[3] == np.float32(1.0 + 2 ** -7)  # 1.5 ulp rounds up
    assert r[4] == -2.5 and r[5] == 0.0


def test_encoding_layout():
    raw = np.zeros((1, 14), np.float32)
    raw[0, :3] = [0.5, -0.25, 0.0]
    raw[0, 3:14] = np.arange(11)
    f = N.encode(raw, np.zeros(3), np.ones(3))
    assert f.shape == (1, 96)
    assert np.isclose(f[0, 0], np.sin(np.pi * 0.5), atol=1e-6) and np.isclose(f[0, 1], np.cos(np.pi * 0.5), atol=1e-6)
    assert np.isclose(f[0, 24 + 2], np.sin(2 * np.pi * -0.25), atol=1e-6)
    assert np.array_equal(f[0, 72:83], np.arange(11))
    assert np.all(f[0, 83:] == 0)


def test_adam_step_formula():
    p = N.init_params(0)
    g = np.random.default_rng(0).normal(size=N.N_PARAMS).astype(np.float32) * 1e-3
    p1, m, v = N.adam(p, np.zeros_like(p), np.zeros_like(p), g, 1, lr=1e-2)
    # the first bias-corrected step moves each trainable weight by lr * g / (|g| + eps)
    trainable = np.array([N._trainable(i) for i in range(N.N_PARAMS)])
    sel = trainable & (np.abs(g) > 1e-6)
    want = 1e-2 * g[sel].astype(np.float64) / (np.abs(g[sel].astype(np.float64)) + 1e-8)
    assert np.allclose((p - p1)[sel], want, rtol=2e-4, atol=2e-8)
    assert np.all(p1[~trainable] == 0)


def test_training_reduces_loss_on_cpu():
    rng = np.random.default_rng(1)
    raw = rng.uniform(-1, 1, (256, 14)).astype(np.float32)
    t = (0.5 + 0.4 * np.sin(2 * raw[:, :3])).astype(np.float32)
    x = N.encode(raw, np.zeros(3), np.ones(3))
    p = N.init_params(2)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    y, _ = N.forward(p, x)
    l0 = N.loss_grad(y, t)[0]
    for step in range(1, 60):
        y, acts = N.forward(p, x)
        _, dy = N.loss_grad(y, t)
        g = N.backward(p, acts, dy)
        p, m, v = N.adam(p, m, v, g, step, lr=5e-3)
    y, _ = N.forward(p, x)
    assert N.loss_grad(y, t)[0] < 0.2 * l0
