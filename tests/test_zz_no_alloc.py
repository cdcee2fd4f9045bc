"""slimrt_render_device_noalloc (the frame entry point that allocates nothing) against slimrt_render_device,
through the C ABI: the same samples bit for bit, call after call, in all three node widths and both tile orders, when
the frame size shrinks between calls (the recorded durations then belong to another tile count), when another
context renders in between (the scratch array is shared per GPU), and with every sample written (canvases filled
with NaN before each call).  Frames beyond the scratch's 131072 tiles and wavefront forms are rejected."""
import ctypes as C

import numpy as np
import pytest

from slimtracincpp_amd import scenes, runtime

pytestmark = pytest.mark.gpu


def _noalloc():
    return runtime.lib().slimrt_render_device_noalloc


def _frame(r, p, noalloc):
    import torch
    n = p.width * p.height * (4 if p.ssaa else 1)
    px = torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda")
    dp = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    if noalloc:
        rc = _noalloc()(r.ctx, C.byref(p), C.c_void_p(px.data_ptr()), C.c_void_p(dp.data_ptr()))
        assert rc == 0, runtime.lib().slimrt_last_error(r.ctx).decode()
    else:
        r.render_device(p, px.data_ptr(), dp.data_ptr())
    r.synchronize()
    torch.cuda.synchronize()
    return px.cpu().numpy().view(np.uint32), dp.cpu().numpy().view(np.uint32)


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def _params(r, size, flags):
    p = r.params(output=runtime.OUTPUT_RADIANCE, size=size)
    p.flags = flags
    return p


@pytest.mark.parametrize("flags", [runtime.BVH4, 0, runtime.BVH4 | runtime.TILE_ORDER_FIXED, runtime.BVH8])
def test_noalloc_equals_render_device(gpu_available, flags):
    spec = scenes.mesh_pbr_scene(0, 640, 360, level=5, texture_size=128, sky_size=32, irradiance_size=8)
    r = runtime.Renderer(spec, "gpu")
    big, small = _params(r, (640, 360), flags), _params(r, (320, 180), flags)
    ref_big = _frame(r, big, noalloc=False)
    ref_small = _frame(r, small, noalloc=False)
    assert not np.any(np.isnan(ref_big[0].view(np.float32)))
    # two calls in a row, the second with a smaller frame; then the larger again, several times (from its second
    # call on, in cost order from the durations its previous call recorded)
    for size_seq in (["big", "small"], ["big", "big", "big", "small", "small", "big"]):
        for which in size_seq:
            p, ref = (big, ref_big) if which == "big" else (small, ref_small)
            assert _same(_frame(r, p, noalloc=True), ref), (flags, which)
    r.close()


def test_noalloc_shared_between_contexts(gpu_available):
    """Two contexts on one GPU share the scratch array: interleaved calls stay bit-identical to render_device."""
    a_spec = scenes.mesh_pbr_scene(1, 480, 270, level=4, texture_size=64, sky_size=16, irradiance_size=8)
    b_spec = scenes.lit_mesh_scene(0, 480, 270, level=4, texture_size=64, sky_size=16, irradiance_size=8)
    a, b = runtime.Renderer(a_spec, "gpu"), runtime.Renderer(b_spec, "gpu")
    pa, pb = _params(a, (480, 270), runtime.BVH4), _params(b, (480, 270), runtime.BVH4)
    ref_a, ref_b = _frame(a, pa, noalloc=False), _frame(b, pb, noalloc=False)
    for _ in range(3):
        assert _same(_frame(a, pa, noalloc=True), ref_a)
        assert _same(_frame(b, pb, noalloc=True), ref_b)
    a.close()
    b.close()


def test_noalloc_rejects_what_does_not_fit(gpu_available):
    spec = scenes.implicit_scene(0, 64, 36)
    r = runtime.Renderer(spec, "gpu")
    fn = _noalloc()
    dummy = C.c_void_p(16)    # never dereferenced: the calls fail before any launch
    p = _params(r, (4096, 2160), 0)                   # 128 x 1080 = 138,240 tiles > 131,072
    assert fn(r.ctx, C.byref(p), dummy, dummy) == -1
    assert "131072" in runtime.lib().slimrt_last_error(r.ctx).decode()
    p = _params(r, (64, 36), runtime.WAVEFRONT)
    assert fn(r.ctx, C.byref(p), dummy, dummy) == -1
    p = _params(r, (3840, 2160), 0)                   # 120 x 1080 = 129,600 tiles: fits
    p.ssaa = 0
    import torch
    px = torch.zeros((3840 * 2160, 4), dtype=torch.float32, device="cuda")
    dp = torch.zeros(3840 * 2160, dtype=torch.float32, device="cuda")
    assert fn(r.ctx, C.byref(p), C.c_void_p(px.data_ptr()), C.c_void_p(dp.data_ptr())) == 0
    r.synchronize()
    ref = _frame(r, p, noalloc=False)
    assert np.array_equal(px.cpu().numpy().view(np.uint32), ref[0])
    r.close()
