"""Bounces as a wavefront (render flag SLIMRT_WAVEFRONT): compacted ray queues run to the paths' ends in one launch
or re-compacted after every bounce, sorted by direction octant or not, in 1, 2, 4 or 8 bands whose bounces
overlap the next band on a second stream, with the first hits traced by the shading kernel or by a separate
traversal kernel (SLIMRT_WAVEFRONT_SPLIT), must give every sample bit for bit what
the per-sample megakernel gives (same iteration code, same accumulation order per sample), on scenes with point
lights and glow, transparency, glass and mirrors, SSAA and both output kinds; and a frame with many bounces
(depth 8) equals the oracle within the beauty tolerance.  With SLIMRT_WAVEFRONT_COUNT the queues' path counters
obey conservation and equal the oracle's per-iteration path counts."""
import numpy as np
import pytest

from slimtracincpp_amd import scenes, runtime
import oracle

pytestmark = pytest.mark.gpu

WAVEFRONT, PER_BOUNCE, SORT, SPLIT = 2, 4, 8, 256


def _frame(r, p):
    import torch
    n = p.width * p.height * (4 if p.ssaa else 1)
    px = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    dp = torch.zeros(n, dtype=torch.float32, device="cuda")
    r.render_device(p, px.data_ptr(), dp.data_ptr())
    r.synchronize()
    torch.cuda.synchronize()
    return px.cpu().numpy().view(np.uint32), dp.cpu().numpy().view(np.uint32)


CASES = {
    "implicit_d2": lambda: scenes.implicit_scene(0, 160, 90),
    "implicit_transparent_d5": lambda: scenes.implicit_scene(1, 120, 72, max_depth=5, transparent=True),
    "mesh_pbr_d4": lambda: scenes.mesh_pbr_scene(0, 128, 72, level=3, texture_size=64, sky_size=16, irradiance_size=8),
    "mesh_pbr_d8": lambda: scenes.mesh_pbr_scene(1, 96, 54, max_depth=8, level=4, texture_size=64, sky_size=16,
                                                 irradiance_size=8),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("flags", [WAVEFRONT, WAVEFRONT | SORT, WAVEFRONT | PER_BOUNCE, WAVEFRONT | PER_BOUNCE | SORT,
                                   WAVEFRONT | (1 << 4), WAVEFRONT | PER_BOUNCE | SORT | (2 << 4),
                                   WAVEFRONT | PER_BOUNCE | (3 << 4), WAVEFRONT | SPLIT,
                                   WAVEFRONT | SORT | SPLIT | (2 << 4), WAVEFRONT | PER_BOUNCE | SPLIT])
@pytest.mark.parametrize("output", [runtime.OUTPUT_CANVAS, runtime.OUTPUT_RADIANCE])
def test_wavefront_equals_megakernel(gpu_available, name, flags, output):
    spec = CASES[name]()
    r = runtime.Renderer(spec, "gpu")
    p = r.params(output=output)
    ref_px, ref_dp = _frame(r, p)
    p.flags = flags
    got_px, got_dp = _frame(r, p)
    assert np.array_equal(got_px, ref_px) and np.array_equal(got_dp, ref_dp)


def test_wavefront_ssaa(gpu_available):
    spec = scenes.implicit_scene(2, 64, 40, max_depth=3, transparent=True)
    spec.ssaa = True
    r = runtime.Renderer(spec, "gpu")
    p = r.params()
    ref = _frame(r, p)
    p.flags = WAVEFRONT
    got = _frame(r, p)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_wavefront_deep_bounces_against_oracle(gpu_available):
    spec = CASES["mesh_pbr_d8"]()
    r = runtime.Renderer(spec, "gpu")
    p = r.params(output=runtime.OUTPUT_RADIANCE)
    p.flags = WAVEFRONT
    px, _ = _frame(r, p)
    col, _, _ = oracle.OracleScene(spec).render()
    got = px.view(np.float32).reshape(spec.height, spec.width, 4)[..., :3]
    assert np.max(np.abs(got - col)) <= 1e-4


def test_render_autotune_picks_a_form_with_the_same_frame(gpu_available):
    spec = CASES["mesh_pbr_d4"]()
    r = runtime.Renderer(spec, "gpu")
    p = r.params()
    import torch
    n = p.width * p.height
    px = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    dp = torch.zeros(n, dtype=torch.float32, device="cuda")
    flags, ms = r.autotune_form(p, px.data_ptr(), dp.data_ptr(), reps=2)
    assert flags & ~(WAVEFRONT | PER_BOUNCE | SORT | (3 << 4) | SPLIT | runtime.BVH4 | runtime.BVH8) == 0 and ms > 0
    assert flags & (PER_BOUNCE | SORT | (3 << 4) | SPLIT) == 0 or flags & WAVEFRONT
    assert not (flags & runtime.BVH8 and flags & WAVEFRONT)   # node width 8 is a megakernel form only
    p.flags = flags & (runtime.BVH4 | runtime.BVH8)   # the megakernel with the same node width: the same bits
    ref = _frame(r, p)
    p.flags = flags
    got = _frame(r, p)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


COUNT_FORMS = [WAVEFRONT, WAVEFRONT | SORT, WAVEFRONT | PER_BOUNCE, WAVEFRONT | PER_BOUNCE | SORT,
               WAVEFRONT | (2 << 4), WAVEFRONT | PER_BOUNCE | (1 << 4), WAVEFRONT | SPLIT | (1 << 4)]


@pytest.mark.parametrize("name", ["implicit_transparent_d5", "mesh_pbr_d8"])
def test_wavefront_path_counts(gpu_available, name):
    """Queue conservation and the per-iteration path counts: in every form each path runs one iteration per queue
    it passes through, every iteration either spawns a bounce ray into a queue or finishes its sample, every
    sample finishes once, and the number of paths running iteration i equals the oracle's count of samples whose
    path reaches iteration i."""
    spec = CASES[name]()
    r = runtime.Renderer(spec, "gpu")
    p = r.params()
    samples = p.width * p.height
    want = oracle.OracleScene(spec).iteration_counts(n=16).astype(np.int64)
    assert want[0] == samples and want[1] > 0            # the case has bounces
    for flags in COUNT_FORMS + [WAVEFRONT | runtime.BVH4, WAVEFRONT | PER_BOUNCE | SORT | runtime.BVH4]:
        p.flags = flags | runtime.WAVEFRONT_COUNT
        _frame(r, p)
        c = r.wavefront_counts().astype(np.int64)
        it, spawned, finished = c[:32], c[32], c[33]
        assert finished == samples, flags
        assert it.sum() == spawned + finished, flags
        assert spawned == it[1:].sum(), flags
        assert np.all(np.diff(it) <= 0), flags
        if not flags & runtime.BVH4:                       # 4-wide walks may resolve exact ties differently
            assert np.array_equal(it[:16], want) and not it[16:].any(), (flags, it[:8], want[:8])
    p.flags = WAVEFRONT                                   # counting leaves the frame's bits alone
    plain = _frame(r, p)
    p.flags = WAVEFRONT | runtime.WAVEFRONT_COUNT
    counted = _frame(r, p)
    assert np.array_equal(plain[0], counted[0]) and np.array_equal(plain[1], counted[1])
