"""4-wide mesh BVHs (render flag SLIMRT_BVH4): collapsed at upload from the sibling-pair BVHs, same leaves and
triangle ranges.  Closest hits are the same except exact ties between triangles, so against the sibling-pair
frames: primary hit ids on >= 99.99% of samples, beauty within 1e-4 per channel (the parity bars), in the
megakernel and in the wavefront forms; debug modes too."""
import numpy as np
import pytest

from slimtracincpp_amd import scenes, runtime

pytestmark = pytest.mark.gpu

BVH4 = 64


def _frame(r, p):
    import torch
    n = p.width * p.height * (4 if p.ssaa else 1)
    px = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    dp = torch.zeros(n, dtype=torch.float32, device="cuda")
    r.render_device(p, px.data_ptr(), dp.data_ptr())
    r.synchronize()
    torch.cuda.synchronize()
    return px.cpu().numpy(), dp.cpu().numpy()


CASES = {
    "mesh_pbr_l6": lambda: scenes.mesh_pbr_scene(0, 256, 144, level=6, texture_size=64, sky_size=16, irradiance_size=8),
    "mesh_pbr_leaf1_d8": lambda: scenes.mesh_pbr_scene(2, 160, 90, max_depth=8, level=4, leaf_size=1, texture_size=64,
                                                       sky_size=16, irradiance_size=8),
    "large_small": lambda: scenes.large_scene(0, 160, 90, uv_sphere=(33, 64), texture_size=64, sky_size=16,
                                              irradiance_size=8),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("form", [0, 2, 2 | 4 | 8, 2 | 256])
def test_bvh4_matches_pair_walk(gpu_available, name, form):
    spec = CASES[name]()
    r = runtime.Renderer(spec, "gpu")
    p = r.params(output=runtime.OUTPUT_RADIANCE)
    p.flags = form
    ref, ref_d = _frame(r, p)
    p.flags = form | BVH4
    got, got_d = _frame(r, p)
    err = np.abs(got[:, :3] - ref[:, :3])
    print(name, form, "max err", err.max(), "bit-exact fraction", np.mean(got.view(np.uint32) == ref.view(np.uint32)))
    assert np.mean(np.all(err <= 1e-4, axis=1)) >= 0.9999
    fin = np.isfinite(ref_d)
    assert np.array_equal(fin, np.isfinite(got_d))


@pytest.mark.parametrize("name", sorted(CASES))
def test_bvh4_primary_hits(gpu_available, name):
    spec = CASES[name]()
    r = runtime.Renderer(spec, "gpu")
    p = r.params(output=runtime.OUTPUT_HITS)
    a, _ = _frame(r, p)
    p.flags = BVH4
    b, _ = _frame(r, p)
    ga, ta = a[:, 0].view(np.int32), a[:, 1].view(np.int32)
    gb, tb = b[:, 0].view(np.int32), b[:, 1].view(np.int32)
    match = np.mean((ga == gb) & (ta == tb))
    print(name, "hit id match", match)
    assert match >= 0.9999
