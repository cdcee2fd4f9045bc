"""Presentation (slimrt_canvas_content): Canvas::drawToWindow's window content of a canvas.

The oracle's restatement equals the reference's own getPixelContent on the reference's canvas (live, through the
compiled reference driver), with and without SSAA.  The library (CPU context, GPU kernel) equals the oracle on
arbitrary canvases, including values above 1, transparent pixels and fully transparent SSAA quads."""
import os

import numpy as np
import pytest

import oracle
from slimtracincpp_amd import runtime, scenes

needs_ref = pytest.mark.skipif(not (os.path.exists(oracle.REF_PATH) and os.path.isdir("/root/reference/src")),
                               reason="compiled reference not available")


def _canvas_of(col, width, ssaa):
    """What Canvas::setPixel (opacity -1) stores for pre-canvas colours (Hs, Ws, 3), in the canvas layout."""
    c = np.clip(col, 0, 1).astype(np.float32)
    v = (c * (c * np.float32(1.0))).astype(np.float32)
    Hs, Ws = v.shape[:2]
    y, x = np.mgrid[0:Hs, 0:Ws]
    o = ((width * (y >> 1) + (x >> 1)) * 4 + 2 * (y & 1) + (x & 1)) if ssaa else (width * y + x)
    out = np.zeros((Hs * Ws, 4), np.float32)
    out[o.reshape(-1), :3] = v.reshape(-1, 3)
    out[:, 3] = 1.0
    return out


@needs_ref
@pytest.mark.parametrize("ssaa", [False, True])
def test_oracle_content_equals_reference(ssaa):
    spec = scenes.mesh_pbr_scene(4, 48, 32, level=2, texture_size=16, sky_size=8, irradiance_size=4)
    spec.ssaa = ssaa
    ref = oracle.run_reference(spec)
    col, _, _ = oracle.OracleScene(spec).render()
    got = oracle.canvas_content(_canvas_of(col, 48, ssaa), 48, 32, ssaa)
    assert np.array_equal(got, ref["content"])
    assert np.count_nonzero(got) > 0.9 * got.size


def _random_canvas(w, h, ssaa, seed):
    rng = np.random.default_rng(seed)
    n = w * h * (4 if ssaa else 1)
    px = rng.random((n, 4)).astype(np.float32) * np.float32(1.3)      # some channels above 1
    px[:, 3] = (rng.random(n) < 0.8).astype(np.float32)                # some transparent samples
    if ssaa:
        px[:16, 3] = 0.0                                                # whole transparent quads
    return px


@pytest.mark.parametrize("ssaa", [False, True])
def test_cpu_context_content_equals_oracle(ssaa):
    spec = scenes.implicit_scene(0, 16, 8)
    r = runtime.Renderer(spec, "cpu")
    px = _random_canvas(40, 24, ssaa, 3)
    out = np.zeros(40 * 24, np.uint32)
    r.canvas_content(px.ctypes.data, 40, 24, ssaa, out.ctypes.data)
    assert np.array_equal(out, oracle.canvas_content(px, 40, 24, ssaa))
    if ssaa:
        assert out[0] == 0 and out[1] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("ssaa", [False, True])
def test_gpu_content(gpu_available, ssaa):
    import torch
    px = _random_canvas(200, 120, ssaa, 5)
    spec = scenes.implicit_scene(1, 64, 40)
    spec.ssaa = ssaa
    r = runtime.Renderer(spec, "gpu")
    d_px = torch.from_numpy(px).cuda()
    d_out = torch.zeros(200 * 120, dtype=torch.int32, device="cuda")
    r.canvas_content(d_px.data_ptr(), 200, 120, ssaa, d_out.data_ptr())
    r.synchronize()
    assert np.array_equal(d_out.cpu().numpy().view(np.uint32), oracle.canvas_content(px, 200, 120, ssaa))
    # a rendered canvas: the frame's own content, and the oracle frame's content within one level per channel
    p = r.params(output=runtime.OUTPUT_CANVAS)
    n = 64 * 40 * (4 if ssaa else 1)
    cpx = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    cdp = torch.zeros(n, dtype=torch.float32, device="cuda")
    r.render_device(p, cpx.data_ptr(), cdp.data_ptr())
    out = torch.zeros(64 * 40, dtype=torch.int32, device="cuda")
    r.canvas_content(cpx.data_ptr(), 64, 40, ssaa, out.data_ptr())
    r.synchronize()
    got = out.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, oracle.canvas_content(cpx.cpu().numpy(), 64, 40, ssaa))
    col, _, _ = oracle.OracleScene(spec).render()
    want = oracle.canvas_content(_canvas_of(col, 64, ssaa), 64, 40, ssaa)
    chan = lambda v, s: ((v >> s) & 0xFF).astype(np.int32)
    assert all(np.max(np.abs(chan(got, s) - chan(want, s))) <= 1 for s in (0, 8, 16))
