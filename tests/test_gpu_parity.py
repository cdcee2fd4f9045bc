"""GPU parity: the gfx950 kernels, called through the C ABI, against the CPU oracle on the same seeded scenes.

Bars (per the parity contract): primary hit ids identical on >= 99.99% of samples; depth, normal and UV debug
modes within 1e-5 relative; beauty radiance within 1e-4 absolute per channel.
"""
import numpy as np
import pytest

from slimtracincpp_amd import scenes, runtime
import oracle

pytestmark = pytest.mark.gpu

BEAUTY_ATOL = 1e-4
DEBUG_RTOL = 1e-5


def _close(a, b, atol=0.0, rtol=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    same_inf = np.array_equal(np.isinf(a), np.isinf(b)) and np.array_equal(np.isnan(a), np.isnan(b))
    fin = np.isfinite(a) & np.isfinite(b)
    err = np.abs(a[fin] - b[fin])
    tol = atol + rtol * np.abs(b[fin])
    return same_inf and bool(np.all(err <= tol)), float(err.max(initial=0.0))


@pytest.fixture(scope="module")
def implicit():
    spec = scenes.implicit_scene(0)
    return spec, runtime.Renderer(spec, "gpu"), oracle.OracleScene(spec)


@pytest.fixture(scope="module")
def implicit_transparent():
    spec = scenes.implicit_scene(1, transparent=True)
    return spec, runtime.Renderer(spec, "gpu"), oracle.OracleScene(spec)


@pytest.mark.parametrize("mode", [scenes.MODE_BEAUTY, scenes.MODE_NORMALS, scenes.MODE_DEPTH, scenes.MODE_UVS,
                                  scenes.MODE_MIP_LEVEL, scenes.MODE_NORMAL_MAP])
def test_implicit_scene_modes(gpu_available, implicit, mode):
    spec, r, o = implicit
    px, depth = r.render(mode=mode)
    col, dep, _ = o.render(mode=mode)
    tol = dict(atol=BEAUTY_ATOL) if mode == scenes.MODE_BEAUTY else dict(rtol=DEBUG_RTOL)
    ok, err = _close(px[..., :3], col, **tol)
    print(f"mode {mode}: max err {err:.3g}, bit-exact {np.array_equal(px[..., :3].view(np.uint32), col.view(np.uint32))}")
    assert ok, err
    ok, err = _close(depth, dep, rtol=DEBUG_RTOL)
    assert ok, err


@pytest.mark.parametrize("mode", [scenes.MODE_BEAUTY, scenes.MODE_NORMALS, scenes.MODE_UVS])
def test_transparency(gpu_available, implicit_transparent, mode):
    spec, r, o = implicit_transparent
    px, depth = r.render(mode=mode)
    col, dep, _ = o.render(mode=mode)
    ok, err = _close(px[..., :3], col, atol=BEAUTY_ATOL if mode == scenes.MODE_BEAUTY else 0, rtol=0 if mode == scenes.MODE_BEAUTY else DEBUG_RTOL)
    assert ok, err


def test_work_counters_match_oracle(gpu_available, implicit):
    spec, r, o = implicit
    _, _, cnt = o.render()
    assert r.count_work() == cnt


def test_canvas_output_is_clamped_square(gpu_available, implicit):
    spec, r, o = implicit
    rad, _ = r.render(output=runtime.OUTPUT_RADIANCE)
    can, _ = r.render(output=runtime.OUTPUT_CANVAS)
    c = np.clip(rad[..., :3], 0, 1)
    assert np.array_equal(can[..., :3], (c * (c * np.float32(1.0))).astype(np.float32))
    assert np.all(can[..., 3] == 1.0)


@pytest.fixture(scope="module")
def mesh_small():
    spec = scenes.mesh_pbr_scene(0, 320, 180, level=4)
    return spec, runtime.Renderer(spec, "gpu"), oracle.OracleScene(spec)


@pytest.mark.parametrize("mode", [scenes.MODE_BEAUTY, scenes.MODE_NORMALS, scenes.MODE_UVS, scenes.MODE_DEPTH])
def test_mesh_pbr_small(gpu_available, mesh_small, mode):
    spec, r, o = mesh_small
    px, depth = r.render(mode=mode)
    col, dep, _ = o.render(mode=mode)
    ok, err = _close(px[..., :3], col, atol=BEAUTY_ATOL if mode == scenes.MODE_BEAUTY else 0, rtol=0 if mode == scenes.MODE_BEAUTY else DEBUG_RTOL)
    print(f"mesh mode {mode}: max err {err:.3g}")
    assert ok, err


def test_shipped_cube_map_sky(gpu_available):
    """Cube-map sampling on the texture the reference's converter wrote (sky colour, radiance, irradiance)."""
    from tests.helpers import shipped_sky_scene
    spec = shipped_sky_scene()
    r, o = runtime.Renderer(spec, "gpu"), oracle.OracleScene(spec)
    px, depth = r.render()
    col, dep, _ = o.render()
    ok, err = _close(px[..., :3], col, atol=BEAUTY_ATOL)
    assert ok, err
    ok, err = _close(depth, dep, rtol=DEBUG_RTOL)
    assert ok, err


@pytest.fixture(scope="module")
def mesh_1080p():
    spec = scenes.mesh_pbr_scene(0, 1920, 1080)
    return spec, runtime.Renderer(spec, "gpu"), oracle.OracleScene(spec)


def test_mesh_pbr_1080p_primary_hits(gpu_available, mesh_1080p):
    """The benchmark workload at full size: primary hit ids (geometry, triangle) on >= 99.99% of samples."""
    spec, r, o = mesh_1080p
    g, t, d = r.primary_hits()
    g2, t2, d2 = o.primary_hits()
    agree = (g == g2) & (t == t2)
    print(f"1080p hit-id agreement {agree.mean():.6f}, mismatches {int((~agree).sum())}")
    assert agree.mean() >= 0.9999
    both = agree & np.isfinite(d) & np.isfinite(d2)
    assert np.all(np.abs(d[both] - d2[both]) <= DEBUG_RTOL * np.abs(d2[both]))


def test_mesh_pbr_1080p_beauty(gpu_available, mesh_1080p):
    spec, r, o = mesh_1080p
    px, depth = r.render()
    col, dep, cnt = o.render()
    ok, err = _close(px[..., :3], col, atol=BEAUTY_ATOL)
    exact = float(np.mean(px[..., :3].view(np.uint32) == col.view(np.uint32)))
    print(f"1080p beauty max err {err:.3g}, bit-exact fraction {exact:.6f}")
    assert ok, err
    ok, err = _close(depth, dep, rtol=DEBUG_RTOL)
    assert ok, err
    assert r.count_work() == cnt


def test_benchmark_form_1080p_against_oracle(gpu_available, mesh_1080p):
    """The frame form the benchmark reports - 4-wide mesh BVHs, tiles costliest first from the fourth frame, the
    costliest tiles as two half waves - at the benchmark's size, directly against the oracle: every frame's
    beauty within 1e-4 and depth within 1e-5 relative, primary hit ids on >= 99.99% of samples."""
    import torch
    spec, r, o = mesh_1080p
    col, dep, _ = o.render()
    p = r.params(output=runtime.OUTPUT_RADIANCE)
    p.flags = runtime.BVH4
    n = p.width * p.height
    for frame in range(5):
        px = torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda")
        dp = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
        r.render_device(p, px.data_ptr(), dp.data_ptr())
        r.synchronize()
        got = px.cpu().numpy().reshape(spec.height, spec.width, 4)
        ok, err = _close(got[..., :3], col, atol=BEAUTY_ATOL)
        exact = float(np.mean(got[..., :3].view(np.uint32) == col.view(np.uint32)))
        print(f"frame {frame}: beauty max err {err:.3g}, bit-exact fraction {exact:.6f}")
        assert ok, (frame, err)
        ok, err = _close(dp.cpu().numpy().reshape(spec.height, spec.width), dep, rtol=DEBUG_RTOL)
        assert ok, (frame, err)
    costs, order = r.tile_costs()
    assert not np.array_equal(order, np.arange(order.size))   # the last frame ran its costliest tiles first
    p.output = runtime.OUTPUT_HITS
    hits, _ = r.render(p)
    g, t, d = hits[..., 0].view(np.int32), hits[..., 1].view(np.int32), hits[..., 2]
    g2, t2, d2 = o.primary_hits()
    agree = (g == g2) & (t == t2)
    print(f"4-wide 1080p hit-id agreement {agree.mean():.6f}, mismatches {int((~agree).sum())}")
    assert agree.mean() >= 0.9999
    both = agree & np.isfinite(d) & np.isfinite(d2)
    assert np.all(np.abs(d[both] - d2[both]) <= DEBUG_RTOL * np.abs(d2[both]))


@pytest.mark.parametrize("mode", [scenes.MODE_NORMALS, scenes.MODE_UVS, scenes.MODE_DEPTH])
def test_mesh_pbr_1080p_debug_modes(gpu_available, mesh_1080p, mode):
    spec, r, o = mesh_1080p
    px, _ = r.render(mode=mode)
    col, _, _ = o.render(mode=mode)
    ok, err = _close(px[..., :3], col, rtol=DEBUG_RTOL)
    assert ok, err


def test_frames_are_deterministic(gpu_available, mesh_small):
    spec, r, o = mesh_small
    a, da = r.render()
    b, db = r.render()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and np.array_equal(da.view(np.uint32), db.view(np.uint32))


def test_ray_cone_mode(gpu_available):
    spec = scenes.mesh_pbr_scene(0, 160, 90, level=3, texture_size=256, sky_size=16, irradiance_size=8)
    r = runtime.Renderer(spec, "gpu")
    o = oracle.OracleScene(spec, ray_cones=True)
    for mode in (scenes.MODE_BEAUTY, scenes.MODE_MIP_LEVEL):
        px, _ = r.render(mode=mode, ray_cones=True)
        col, _, _ = o.render(mode=mode)
        ok, err = _close(px[..., :3], col, atol=BEAUTY_ATOL)
        assert ok, (mode, err)


def test_ssaa_on_gpu(gpu_available):
    spec = scenes.implicit_scene(0, 64, 36)
    spec.ssaa = True
    r = runtime.Renderer(spec, "gpu")
    c = runtime.Renderer(spec, "cpu")
    a, _ = r.render()
    b, _ = c.render()
    ok, err = _close(a, b, atol=BEAUTY_ATOL)
    assert ok, err


def test_stack_overflow_reported_on_gpu(gpu_available):
    spec = scenes.mesh_pbr_scene(0, 32, 18, level=3, texture_size=16, sky_size=8, irradiance_size=4)
    sb = runtime.build_scene(spec)
    sb.meshes[0].height = 1
    sb.mesh_descs[0].height = 1
    r = runtime.Renderer(sb, "gpu")
    with pytest.raises(runtime.SlimrtError) as e:
        r.render()
    assert e.value.code == -5


def test_moving_geometries_rebuild(gpu_available):
    """Per-frame transform updates (Scene::updateAABBs + updateBVH, then uploadGeometries + uploadSceneBVH)."""
    spec = scenes.implicit_scene(0, 96, 54)
    r = runtime.Renderer(spec, "gpu")
    rng = np.random.default_rng(3)
    for frame in range(3):
        gf = spec.geo_f.copy()
        gf[1:, 3:6] += rng.uniform(-0.5, 0.5, (gf.shape[0] - 1, 3)).astype(np.float32)
        gf[1:, 0:3] += np.float32(0.3 * frame)
        r.update_geometries(gf)
        moved = scenes.SceneSpec(**{**spec.__dict__, "geo_f": gf})
        col, _, _ = oracle.OracleScene(moved).render()
        px, _ = r.render()
        ok, err = _close(px[..., :3], col, atol=BEAUTY_ATOL)
        assert ok, (frame, err)


@pytest.mark.parametrize("golden", ["large_s0_160x90_emissive"])
def test_golden_reference_frames_on_gpu(gpu_available, golden, tmp_path):
    """The GPU against the reference's own frame (golden vectors written by the compiled reference): the large
    workload's generator at a small size - 53 geometries, emissive quads, normal maps, glass, mirrors, depth 8."""
    from tests.helpers import load_golden, golden_spec_path
    from tests.golden.make_golden import CASES
    g = load_golden(golden)
    golden_spec_path(golden, g, str(tmp_path))   # checks the seeded generator against the golden's digest
    spec = CASES[golden]()
    r = runtime.Renderer(spec, "gpu")
    px, depth = r.render()
    col = g["color"].reshape(spec.height, spec.width, 3)
    ok, err = _close(px[..., :3], col, atol=BEAUTY_ATOL)
    assert ok, err
    geo, tri, t = r.primary_hits()
    match = np.mean((geo.reshape(-1) == g["hit_geo"]) & (tri.reshape(-1) == g["hit_tri"]))
    assert match >= 0.9999, match
    p = r.params()
    p.flags = runtime.WAVEFRONT
    pw, dw = r.render(p)
    assert np.array_equal(pw.view(np.uint32), px.view(np.uint32)) and np.array_equal(dw.view(np.uint32), depth.view(np.uint32))
