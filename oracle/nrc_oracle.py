"""CPU oracle for the neural radiance cache (test infrastructure, not product code).

The cache has no counterpart in the reference; this is a plain restatement of the network the HIP kernels
implement (slimtracincpp_amd/csrc/nrc.hip), used to check their forward pass, gradients and Adam step:

  encoding   position p normalised to the scene box, 12 frequency bands per axis (72 values): sin and cos of
             2^b pi p, computed as one turn times frac(2^(b-1) p) (the reduction is exact in float32), then
             normal (3), view direction (3), roughness, metalness, albedo (3): 83 values, zero-padded to 96
  network    96 -> 64 -> 64 -> 64 -> 64 -> 3 (padded to 16), ReLU after each hidden layer, linear output
  numerics   inputs, weights and hidden activations rounded to bfloat16 (round to nearest even), products
             accumulated in float32 - the MFMA contract; output, loss and gradients in float32
  loss       relative L2 (Mueller et al., "Real-time Neural Radiance Caching for Path Tracing", 2021):
             mean over samples and channels of (y - t)^2 / (stop_grad(y)^2 + 0.01)
  optimiser  Adam (Kingma & Ba, 2015), beta1 0.9, beta2 0.99, eps 1e-8

Parameter layout (float32, 19,728 values): W0 (96x64) b0 (64) W1 (64x64) b1 W2 b2 W3 b3 W4 (64x16) b4 (16);
W[i][o] multiplies input i into output o; W4 columns 3..15 and b4[3..15] stay zero.

Pinning: `check_gradients` compares the analytic backward pass (without bf16 rounding, in float64) with central
finite differences of the loss; the bf16 emulation is pinned against the GPU kernels by the parity tests.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HIDDEN, N_OUT, N_OUT_PAD, N_BANDS = 96, 64, 3, 16, 12
N_ENCODED = 3 * N_BANDS * 2 + 11   # 83
LAYERS = [(N_IN, N_HIDDEN), (N_HIDDEN, N_HIDDEN), (N_HIDDEN, N_HIDDEN), (N_HIDDEN, N_HIDDEN), (N_HIDDEN, N_OUT_PAD)]


def param_offsets():
    offs, o = [], 0
    for fi, fo in LAYERS:
        offs.append((o, o + fi * fo, fi, fo))
        o += fi * fo + fo
    return offs, o


OFFSETS, N_PARAMS = param_offsets()


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest even) and return it as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    return np.where(np.isnan(x), x, out)


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases; padded output columns zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    for li, (wo, bo, fi, fo) in enumerate(OFFSETS):
        lim = np.sqrt(6.0 / fi)
        w = rng.uniform(-lim, lim, (fi, fo)).astype(np.float32)
        if li == len(OFFSETS) - 1:
            w[:, N_OUT:] = 0
        p[wo:wo + fi * fo] = w.reshape(-1)
    return p


def unpack(p: np.ndarray):
    return [(p[wo:wo + fi * fo].reshape(fi, fo), p[wo + fi * fo:wo + fi * fo + fo]) for wo, _, fi, fo in OFFSETS]


def encode(raw: np.ndarray, center: np.ndarray, inv_half: np.ndarray) -> np.ndarray:
    """raw (B, 14): position xyz, normal xyz, view xyz, roughness, metalness, albedo rgb -> (B, 96) float32.
    The position is mapped to [-1, 1]^3 by (p - center) * inv_half before the frequency bands."""
    raw = np.asarray(raw, np.float32)
    B = raw.shape[0]
    f = np.zeros((B, N_IN), np.float32)
    p = ((raw[:, 0:3] - np.asarray(center, np.float32)) * np.asarray(inv_half, np.float32)).astype(np.float32)
    for d in range(3):
        for b in range(N_BANDS):
            x = (p[:, d] * np.float32(2.0 ** (b - 1))).astype(np.float32)     # exact
            turns = (x - np.floor(x)).astype(np.float32)                       # exact
            arg = 2.0 * np.pi * turns.astype(np.float64)
            f[:, 24 * d + 2 * b] = np.sin(arg).astype(np.float32)
            f[:, 24 * d + 2 * b + 1] = np.cos(arg).astype(np.float32)
    f[:, 72:83] = raw[:, 3:14]
    return f


def forward(params: np.ndarray, x: np.ndarray, emulate_bf16: bool = True):
    """Returns the output (B, 3) and the per-layer (input, pre-activation) list for backward."""
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    acts = []
    h = q(x)
    layers = unpack(params)
    for li, (w, b) in enumerate(layers):
        z = (h.astype(np.float64) @ q(w).astype(np.float64)).astype(np.float32 if emulate_bf16 else np.float64) + b
        acts.append((h, z))
        if li < len(layers) - 1:
            h = q(np.maximum(z, 0))
    return z[:, :N_OUT], acts


def loss_grad(y: np.ndarray, t: np.ndarray):
    """Relative L2 loss and its gradient with respect to y (stop-gradient in the denominator)."""
    y64, t64 = y.astype(np.float64), t.astype(np.float64)
    den = y64 * y64 + 0.01
    n = y.size
    return float(np.sum((y64 - t64) ** 2 / den) / n), (2.0 * (y64 - t64) / den / n)


def backward(params: np.ndarray, acts, dy: np.ndarray, emulate_bf16: bool = True) -> np.ndarray:
    """Gradient of the loss with respect to the flat parameter vector."""
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params)
    g = np.zeros(N_PARAMS, np.float64)
    B = dy.shape[0]
    dz = np.zeros((B, N_OUT_PAD))
    dz[:, :N_OUT] = dy
    for li in range(len(layers) - 1, -1, -1):
        w, _ = layers[li]
        h, z = acts[li]
        wo, bo, fi, fo = OFFSETS[li]
        dzq = q(dz.astype(np.float32)) if emulate_bf16 else dz
        g[wo:wo + fi * fo] = (np.asarray(h, np.float64).T @ np.asarray(dzq, np.float64)).reshape(-1)
        g[wo + fi * fo:wo + fi * fo + fo] = dz.sum(0)
        if li == 0:
            break
        dh = np.asarray(dzq, np.float64) @ np.asarray(q(w), np.float64).T
        _, zprev = acts[li - 1]
        dz = dh * (np.asarray(zprev) > 0)
        if emulate_bf16:
            dz = dz.astype(np.float32)
    if emulate_bf16:
        wo, _, fi, fo = OFFSETS[-1]
        g4 = g[wo:wo + fi * fo].reshape(fi, fo)
        g4[:, N_OUT:] = 0
        g[OFFSETS[-1][1] + N_OUT:OFFSETS[-1][1] + N_OUT_PAD] = 0
    return g


def adam(params, m, v, g, step, lr=5e-3, b1=0.9, b2=0.99, eps=1e-8):
    """One Adam step in float32, as the kernel computes it; returns new (params, m, v)."""
    g = g.astype(np.float32)
    m = (np.float32(b1) * m + np.float32(1 - b1) * g).astype(np.float32)
    v = (np.float32(b2) * v + np.float32(1 - b2) * g * g).astype(np.float32)
    mh = m / np.float32(1 - b1 ** step)
    vh = v / np.float32(1 - b2 ** step)
    p = (params - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))).astype(np.float32)
    wo, _, fi, fo = OFFSETS[-1]
    p[wo:wo + fi * fo].reshape(fi, fo)[:, N_OUT:] = 0
    p[OFFSETS[-1][1] + N_OUT:OFFSETS[-1][1] + N_OUT_PAD] = 0
    return p, m, v


def check_gradients(seed: int = 0, B: int = 8, eps: float = 1e-4, n_checks: int = 40) -> float:
    """Max relative error between analytic float64 gradients and central differences of the float64 loss."""
    rng = np.random.default_rng(seed)
    p = init_params(seed).astype(np.float64)
    x = rng.normal(size=(B, N_IN))
    t = rng.uniform(0, 2, (B, N_OUT))

    def loss(pp):
        y, _ = forward(pp, x, emulate_bf16=False)
        return loss_grad_fixed(y, t, den)[0]

    y, acts = forward(p, x, emulate_bf16=False)
    den = y * y + 0.01          # the stop-gradient denominator is held fixed for the difference quotient

    def loss_grad_fixed(yy, tt, dd):
        return float(np.sum((yy - tt) ** 2 / dd) / yy.size), None

    _, dy = loss_grad(y, t)
    g = backward(p, acts, dy, emulate_bf16=False)
    idx = rng.choice([i for i in range(N_PARAMS) if _trainable(i)], n_checks, replace=False)
    worst = 0.0
    for i in idx:
        pp = p.copy()
        pp[i] += eps
        lp = loss(pp)
        pp[i] -= 2 * eps
        lm = loss(pp)
        fd = (lp - lm) / (2 * eps)
        worst = max(worst, abs(fd - g[i]) / max(1e-6, abs(fd) + abs(g[i])))
    return worst


def _trainable(i: int) -> bool:
    wo, bo, fi, fo = OFFSETS[-1]
    if wo <= i < wo + fi * fo:
        return (i - wo) % fo < N_OUT
    if wo + fi * fo <= i < wo + fi * fo + fo:
        return i - (wo + fi * fo) < N_OUT
    return True
