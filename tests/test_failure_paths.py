"""Failure reporting (no counterpart in the reference, whose gpuErrchk aborts or does nothing,
src/slim/core/base.h:20-26):

* NaN samples: a scene whose inputs make some samples NaN (a NaN albedo; the reference passes it through to the
  pixel) sets SLIMRT_FLAG_NAN_SAMPLE in slimrt_frame_flags, in the CPU context and in every GPU frame form; a
  clean frame leaves the flags 0; the flags clear when read; the frame itself succeeds.
* Stack overflow: SLIMRT_FLAG_STACK_OVERFLOW beside the SLIMRT_E_OVERFLOW error (tests/test_host_api.py and
  test_gpu_parity.py force it with a declared BVH height that is too small).
* Allocation failure: with a device memory limit (slimrt_set_memory_limit) an upload or a frame that needs more
  fails with SLIMRT_E_MEMORY, the context keeps what it holds and works again once the limit is lifted.
"""
import numpy as np
import pytest

import oracle
from slimtracincpp_amd import runtime, scenes


def _nan_scene(w=64, h=36):
    spec = scenes.implicit_scene(0, w, h)
    spec.mat_f = spec.mat_f.copy()
    spec.mat_f[3, 0] = np.nan          # one sphere's albedo (red channel)
    return spec


def test_nan_samples_flagged_on_cpu():
    spec = _nan_scene()
    r = runtime.Renderer(spec, "cpu")
    px, dp = r.render()
    col, _, _ = oracle.OracleScene(spec).render()
    nan = np.isnan(px[..., :3]).any(-1)
    assert nan.any() and np.array_equal(nan, np.isnan(col).any(-1))   # the reference's NaN, passed through
    assert r.frame_flags() == runtime.FLAG_NAN_SAMPLE
    assert r.frame_flags() == 0                                       # cleared when read
    clean = runtime.Renderer(scenes.implicit_scene(0, 64, 36), "cpu")
    clean.render()
    assert clean.frame_flags() == 0


def test_memory_functions_on_cpu_context():
    r = runtime.Renderer(scenes.implicit_scene(0, 16, 16), "cpu")
    r.set_memory_limit(1 << 20)
    assert r.memory_usage() == (0, 1 << 20)       # a CPU context holds no device memory


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, runtime.BVH4, runtime.TILE_ORDER_FIXED, runtime.WAVEFRONT,
                                   runtime.WAVEFRONT | runtime.WAVEFRONT_SPLIT | runtime.BVH4,
                                   runtime.WAVEFRONT | runtime.WAVEFRONT_PER_BOUNCE])
def test_nan_samples_flagged_on_gpu(gpu_available, flags):
    spec = _nan_scene(96, 54)
    r = runtime.Renderer(spec, "gpu")
    p = r.params(output=runtime.OUTPUT_RADIANCE)
    p.flags = flags
    px, _ = r.render(p)
    assert np.isnan(px[..., :3]).any()
    assert r.frame_flags() == runtime.FLAG_NAN_SAMPLE
    assert r.frame_flags() == 0
    p.output = runtime.OUTPUT_HITS              # hit records are not samples: no flag
    r.render(p)
    assert r.frame_flags() == 0
    clean = runtime.Renderer(scenes.implicit_scene(0, 96, 54), "gpu")
    q = clean.params()
    q.flags = flags
    clean.render(q)
    assert clean.frame_flags() == 0


@pytest.mark.gpu
def test_stack_overflow_flag_on_gpu(gpu_available):
    spec = scenes.mesh_pbr_scene(0, 32, 18, level=3, texture_size=16, sky_size=8, irradiance_size=4)
    sb = runtime.build_scene(spec)
    sb.meshes[0].height = 1
    sb.mesh_descs[0].height = 1
    r = runtime.Renderer(sb, "gpu")
    p = r.params()
    import torch
    n = p.width * p.height
    px = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    dp = torch.zeros(n, dtype=torch.float32, device="cuda")
    r.render_device(p, px.data_ptr(), dp.data_ptr())
    assert r.frame_flags() & runtime.FLAG_STACK_OVERFLOW


@pytest.mark.gpu
def test_memory_limit_gives_e_memory(gpu_available):
    spec = scenes.mesh_pbr_scene(0, 320, 180, level=4, texture_size=256, sky_size=32, irradiance_size=8)
    sb = runtime.build_scene(spec)
    r = runtime.Renderer(scenes.implicit_scene(0, 16, 16), "gpu")     # a context holding a small scene
    used0, _ = r.memory_usage()
    assert used0 > 0
    r.set_memory_limit(used0 + 4096)
    d = sb.desc()
    rc = runtime.lib().slimrt_upload_scene(r.ctx, runtime.C.byref(d))
    assert rc == -3, rc
    assert "memory" in runtime.lib().slimrt_last_error(r.ctx).decode().lower()
    assert r.memory_usage()[0] <= used0 + 4096
    with pytest.raises(runtime.SlimrtError) as e:                   # a failed upload leaves no scene behind
        r.render()
    assert e.value.code == -4
    # lifted: the upload and a frame work; then a limit below what a larger canvas needs
    r.set_memory_limit(0)
    r.scene = sb
    runtime._check(runtime.lib().slimrt_upload_scene(r.ctx, runtime.C.byref(d)), r.ctx)
    px, _ = r.render()
    assert np.isfinite(px).all()
    used1, _ = r.memory_usage()
    r.set_memory_limit(used1 + 1024)
    with pytest.raises(runtime.SlimrtError) as e:
        r.render(r.params(size=(1920, 1080)))                       # needs a 41 MB canvas on the device
    assert e.value.code == -3
    px2, _ = r.render()                                              # the held canvas still serves its size
    assert np.array_equal(px2.view(np.uint32), px.view(np.uint32))
    r.set_memory_limit(0)
    big, _ = r.render(r.params(size=(1920, 1080)))
    assert big.shape == (1080, 1920, 4)
