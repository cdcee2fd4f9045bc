"""Neural radiance cache kernels (BF16 MFMA, FP32 accumulation) against the CPU oracle (oracle/nrc_oracle.py),
which emulates the same bf16 rounding points.  Tolerances: encodings use the GPU's sinf/cosf against the
host's, and a 1-ulp difference can flip a bf16 rounding, so outputs and gradients are compared by relative norm."""
import numpy as np
import pytest

from oracle import nrc_oracle as N

pytestmark = pytest.mark.gpu

CENTER = np.array([0.5, 1.0, -0.25], np.float32)
INV_HALF = np.array([0.2, 0.25, 0.2], np.float32)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, 14), np.float32)
    raw[:, 0:3] = CENTER + rng.uniform(-4, 4, (n, 3))
    nrm = rng.normal(size=(n, 3))
    raw[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    vw = rng.normal(size=(n, 3))
    raw[:, 6:9] = vw / np.linalg.norm(vw, axis=1, keepdims=True)
    raw[:, 9:11] = rng.uniform(0, 1, (n, 2))
    raw[:, 11:14] = rng.uniform(0, 1, (n, 3))
    t = (0.3 + 0.5 * np.abs(np.sin(raw[:, 0:3] * 0.7) * np.cos(raw[:, 1:2]))).astype(np.float32)
    return raw, t


@pytest.fixture
def cache(gpu_available):
    from slimtracincpp_amd.nrc import NeuralCache
    c = NeuralCache(CENTER, INV_HALF, seed=7)
    c.set_params(N.init_params(5))
    return c


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_inference_matches_oracle(cache):
    import torch
    raw, _ = _data(1000, 1)          # 1000 = 62 full tiles + a partial one
    out = torch.zeros((1000, 3), dtype=torch.float32, device="cuda")
    cache.infer(_cuda(raw), out)
    torch.cuda.synchronize()
    y_gpu = out.cpu().numpy()
    y_ref, _ = N.forward(cache.params(), N.encode(raw, CENTER, INV_HALF))
    rel = np.linalg.norm(y_gpu - y_ref) / np.linalg.norm(y_ref)
    print(f"inference relative error {rel:.2e}, max abs {np.abs(y_gpu - y_ref).max():.2e}")
    assert rel < 2e-3
    assert np.median(np.abs(y_gpu - y_ref)) < 1e-3


def test_gradients_match_oracle(cache):
    raw, t = _data(1000, 2)
    g_gpu, loss_gpu = cache.gradients(_cuda(raw), _cuda(t))
    p = cache.params()
    y, acts = N.forward(p, N.encode(raw, CENTER, INV_HALF))
    loss_ref, dy = N.loss_grad(y, t)
    g_ref = N.backward(p, acts, dy)
    assert abs(loss_gpu - loss_ref) <= 1e-3 * loss_ref
    for li, (wo, bo, fi, fo) in enumerate(N.OFFSETS):
        for lo, hi in ((wo, wo + fi * fo), (wo + fi * fo, wo + fi * fo + fo)):
            a, b = g_gpu[lo:hi], g_ref[lo:hi]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
            cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
            print(f"layer {li} [{lo}:{hi}] relative error {rel:.2e} cosine {cos:.6f}")
            assert rel < 2e-2 and cos > 0.9995


def test_adam_step_matches_oracle(cache):
    raw, t = _data(512, 3)
    p0 = cache.params()
    g, _ = cache.gradients(_cuda(raw), _cuda(t))
    cache.train_step(_cuda(raw), _cuda(t), lr=1e-2, want_loss=True)
    p1 = cache.params()
    want, _, _ = N.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, lr=1e-2)
    assert np.max(np.abs(p1 - want)) <= 1e-6


def test_gradients_deterministic(cache):
    raw, t = _data(700, 4)
    a, la = cache.gradients(_cuda(raw), _cuda(t))
    b, lb = cache.gradients(_cuda(raw), _cuda(t))
    assert np.array_equal(a, b) and la == lb


def test_training_converges(cache):
    raw, t = _data(4096, 5)
    r, tt = _cuda(raw), _cuda(t)
    first = cache.train_step(r, tt, lr=5e-3, want_loss=True)
    for _ in range(300):
        cache.train_step(r, tt, lr=5e-3)
    last = cache.train_step(r, tt, lr=5e-3, want_loss=True)
    print(f"loss {first:.4f} -> {last:.5f}")
    assert last < 0.1 * first


def test_training_state_round_trip(cache):
    """A checkpoint (params, Adam moments, step) restored into another cache continues training bit for bit."""
    from slimtracincpp_amd.nrc import NeuralCache
    raw, t = _data(4096, 11)
    raw_d, t_d = _cuda(raw), _cuda(t)
    for _ in range(3):
        cache.train_step(raw_d, t_d, lr=5e-3)
    saved = cache.state()
    assert saved["step"] == 3 and np.any(saved["m"] != 0) and np.any(saved["v"] != 0)
    for _ in range(2):
        cache.train_step(raw_d, t_d, lr=5e-3)
    after = cache.state()
    other = NeuralCache(CENTER, INV_HALF, seed=99)
    other.set_state(saved)
    back = other.state()
    assert all(np.array_equal(back[k].view(np.uint32), saved[k].view(np.uint32)) for k in ("params", "m", "v"))
    for _ in range(2):
        other.train_step(raw_d, t_d, lr=5e-3)
    got = other.state()
    assert got["step"] == after["step"] == 5
    for k in ("params", "m", "v"):
        assert np.array_equal(got[k].view(np.uint32), after[k].view(np.uint32)), k
    other.close()
