"""Rendering with the neural radiance cache (slimrt_render_nrc).

CPU (oracle): the per-vertex training targets of a path, L_i = current_i + next_i * L_{i+1}, must give back the
pixel: tone_map(L_1) is the beauty colour the integrator writes (bit-exact for one-iteration paths, where both
sides compute the same sums; within 1e-6 otherwise, where the forward and backward sums round differently).

GPU: (1) with no cache vertex reachable, or with the confidence gate closed, a frame equals the plain render bit
for bit, training waves included; (2) the frame's training samples (hit record, target, pixel, vertex) match the
oracle's per-path samples within the beauty tolerance (1e-4, libm differences); (3) training on frames makes
the cached image converge to the traced one: PSNR against the plain render rises by more than 10 dB over the
untrained cache and passes 25 dB; with the gate on, the frame is closer still.
"""
import numpy as np
import pytest

from slimtracincpp_amd import scenes, runtime
import oracle

BEAUTY_ATOL = 1e-4


def _scenes():
    return {
        "implicit_d3": scenes.implicit_scene(1, 160, 90, max_depth=3, transparent=True),
        "mesh_pbr": scenes.mesh_pbr_scene(0, 160, 90, level=3, texture_size=64, sky_size=16, irradiance_size=8),
    }


@pytest.mark.parametrize("name", ["implicit_d3", "mesh_pbr"])
def test_oracle_path_targets_give_back_pixel(name):
    spec = _scenes()[name]
    o = oracle.OracleScene(spec)
    col, _, _ = o.render()
    checked = multi = 0
    for k in range(0, o.samples_w * o.height, 7):
        x, y = k % o.samples_w, k // o.samples_w
        raw, L, v = o.path_samples(x, y)
        if len(v) == 0:
            continue
        assert list(v) == list(range(1, len(v) + 1)), "shaded hits are consecutive path vertices"
        mapped = np.array([oracle.lib().oracle_tone_map(float(c)) for c in L[0]], np.float32)
        if len(v) == 1:
            assert np.array_equal(mapped, col[y, x]), (x, y, mapped, col[y, x])
        else:
            # the later vertices contribute: dropping them (or their throughput) would move the pixel
            tail = np.array([oracle.lib().oracle_tone_map(float(c)) for c in (L[0] - L[1])], np.float32)
            multi += int(np.max(np.abs(tail - col[y, x])) > 1e-3)
            np.testing.assert_allclose(mapped, col[y, x], rtol=0, atol=1e-6)
        # the hit record's view direction is a unit vector and its normal too
        assert abs(np.linalg.norm(raw[0, 6:9]) - 1) < 1e-5 and abs(np.linalg.norm(raw[0, 3:6]) - 1) < 1e-5
        checked += 1
    assert checked > 1000 and multi > 10, (checked, multi)


# ---------------------------------------------------------------- GPU
def _frame_setup(spec):
    import torch
    r = runtime.Renderer(spec, "gpu")
    s = torch.cuda.Stream()
    r.set_stream(s.cuda_stream)
    p = r.params(mode=scenes.MODE_BEAUTY, output=runtime.OUTPUT_RADIANCE)
    n = p.width * p.height
    px = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    dp = torch.zeros(n, dtype=torch.float32, device="cuda")
    return r, s, p, px, dp


def _plain(r, s, p, px, dp):
    import torch
    r.render_device(p, px.data_ptr(), dp.data_ptr())
    s.synchronize()
    r.synchronize()
    return px.cpu().numpy().copy(), dp.cpu().numpy().copy()


def _cached(frames, s, p, px, dp, **kw):
    frames.render_device(p, px.data_ptr(), dp.data_ptr(), **kw)
    s.synchronize()
    frames.r.synchronize()
    return px.cpu().numpy().copy(), dp.cpu().numpy().copy()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["implicit_d3", "mesh_pbr"])
def test_frame_without_cache_vertex_equals_render(gpu_available, name):
    from slimtracincpp_amd.nrc import CachedFrames
    spec = _scenes()[name]
    r, s, p, px, dp = _frame_setup(spec)
    ref_px, ref_dp = _plain(r, s, p, px, dp)
    px.zero_()
    frames = CachedFrames(r, cache_depth=p.max_depth, train_one_in=2, train_steps=1, train_batch=1024,
                          gate_threshold=0.0)
    got_px, got_dp = _cached(frames, s, p, px, dp)
    st = frames.stats()
    print(name, st)
    assert np.array_equal(got_px.view(np.uint32), ref_px.view(np.uint32))
    assert np.array_equal(got_dp.view(np.uint32), ref_dp.view(np.uint32))
    assert st["queries"] == 0 and st["fallbacks"] == 0 and st["samples"] > 0 and st["train_pixels"] > 0
    assert np.isfinite(st["loss_first"]) and st["loss_first"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, runtime.BVH4, runtime.BVH8])   # node width 2, 4, 8
def test_frame_with_gate_closed_equals_render(gpu_available, flags):
    from slimtracincpp_amd.nrc import CachedFrames
    spec = _scenes()["mesh_pbr"]
    r, s, p, px, dp = _frame_setup(spec)
    p.flags = flags
    ref_px, ref_dp = _plain(r, s, p, px, dp)
    px.zero_()
    frames = CachedFrames(r, cache_depth=0, train_one_in=0, gate_threshold=1.0, gate_min_frames=1 << 30)
    got_px, got_dp = _cached(frames, s, p, px, dp)
    st = frames.stats()
    assert np.array_equal(got_px.view(np.uint32), ref_px.view(np.uint32))
    assert np.array_equal(got_dp.view(np.uint32), ref_dp.view(np.uint32))
    # every sample with a shaded primary hit reached the cache vertex and traced on
    hits = int(np.sum(np.isfinite(ref_dp)))
    assert st["queries"] == 0 and st["fallbacks"] == hits, (st, hits)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["implicit_d3", "mesh_pbr"])
def test_training_samples_match_oracle(gpu_available, name):
    from slimtracincpp_amd.nrc import CachedFrames
    spec = _scenes()[name]
    r, s, p, px, dp = _frame_setup(spec)
    frames = CachedFrames(r, cache_depth=p.max_depth, train_one_in=1, train_steps=0, gate_threshold=0.0)
    _cached(frames, s, p, px, dp)
    got = frames.samples()
    st = frames.stats()
    assert got["count"] == st["samples"] and st["train_pixels"] == p.width * p.height
    o = oracle.OracleScene(spec)
    W = p.width
    by_pixel = {}
    for i in range(got["count"]):
        by_pixel.setdefault(int(got["pixel"][i]), []).append(i)
    hits = int(np.sum(np.isfinite(_plain(r, s, p, px, dp)[1])))
    assert len(by_pixel) == hits
    rng = np.random.default_rng(0)
    keys = sorted(by_pixel)
    worst = 0.0
    for k in rng.choice(keys, size=min(400, len(keys)), replace=False):
        idx = sorted(by_pixel[int(k)], key=lambda i: got["vertex"][i])
        raw, L, v = o.path_samples(int(k) % W, int(k) // W)
        assert [int(got["vertex"][i]) for i in idx] == list(v)
        np.testing.assert_allclose(got["raw"][idx], raw, rtol=1e-5, atol=1e-5)
        err = np.abs(got["target"][idx] - L) / np.maximum(1.0, np.abs(L))
        worst = max(worst, float(err.max()))
    print(name, "samples", got["count"], "worst target error", worst)
    assert worst <= BEAUTY_ATOL


def _psnr(a, b):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return 10 * np.log10(1.0 / max(mse, 1e-20))


@pytest.mark.gpu
def test_cache_learns_the_scene(gpu_available):
    """Cache at the primary hit (every pixel answered by the network): the trained cache beats the untrained one by
    a wide margin; the gate then sends the cells it is unsure of back to the tracer."""
    from slimtracincpp_amd.nrc import CachedFrames
    spec = scenes.mesh_pbr_scene(0, 320, 180, level=4, texture_size=128, sky_size=32, irradiance_size=8)
    r, s, p, px, dp = _frame_setup(spec)
    ref, _ = _plain(r, s, p, px, dp)
    frames = CachedFrames(r, cache_depth=0, train_one_in=4, train_steps=4, train_batch=8192, learning_rate=5e-3,
                          gate_threshold=0.0)
    untrained, _ = _cached(frames, s, p, px, dp, train_steps=0, train_one_in=0)
    for _ in range(150):
        _cached(frames, s, p, px, dp)
    st = frames.stats()
    trained, _ = _cached(frames, s, p, px, dp, train_one_in=0)
    st_q = frames.stats()
    p0, p1 = _psnr(untrained[:, :3], ref[:, :3]), _psnr(trained[:, :3], ref[:, :3])
    mae = float(np.mean(np.abs(trained[:, :3] - ref[:, :3])))
    print(f"untrained PSNR {p0:.2f} dB, trained {p1:.2f} dB, MAE {mae:.4f}, last loss {st['loss_last']:.4f}, {st_q}")
    assert st_q["queries"] == int(np.sum(np.isfinite(_plain(r, s, p, px, dp)[1])))
    assert p1 > p0 + 10 and p1 > 25, (p0, p1)
    # gate on: cells whose targets vary more than 10% trace instead
    for _ in range(8):
        _cached(frames, s, p, px, dp, gate_threshold=0.1, gate_min_frames=4)
    gated, _ = _cached(frames, s, p, px, dp, gate_threshold=0.1, gate_min_frames=4, train_one_in=0)
    sg = frames.stats()
    pg = _psnr(gated[:, :3], ref[:, :3])
    print(f"gated PSNR {pg:.2f} dB, {sg}")
    assert sg["fallbacks"] > 0 and sg["queries"] > 0 and pg >= p1
