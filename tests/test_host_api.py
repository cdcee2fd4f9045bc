"""The C ABI library on the host: it loads and exports every declared symbol, its host-side scene derivation equals
the oracle's bit for bit, its CPU context (the CPU side of RayTracingRenderer's use_GPU toggle) renders exactly
what the oracle renders, and its error paths report the documented codes.  No GPU compute runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from slimtracincpp_amd import runtime, scenes
from tests.helpers import bits_equal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(slimrt_[a-z_0-9]+)\s*\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = runtime.lib()
    names = declared_symbols()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n


SCENES = {
    "implicit": lambda: scenes.implicit_scene(0, 96, 54),
    "implicit_t": lambda: scenes.implicit_scene(2, 96, 54, max_depth=3, transparent=True),
    "mesh": lambda: scenes.mesh_pbr_scene(0, 96, 54, level=3, texture_size=64, sky_size=16, irradiance_size=8),
    "lit_mesh": lambda: scenes.lit_mesh_scene(1, 96, 54, level=3, texture_size=32, sky_size=8, irradiance_size=4),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_host_derivation_equals_oracle(name):
    spec = SCENES[name]()
    sb = runtime.build_scene(spec)
    o = oracle.OracleScene(spec)
    assert bits_equal(sb.geometries["orientation"].reshape(-1), o.get("geo_quat"))
    assert bits_equal(sb.aabbs.reshape(-1), o.get("geo_aabb"))
    f = np.concatenate([sb.scene_nodes["min"], sb.scene_nodes["max"]], 1).reshape(-1)
    assert bits_equal(f, o.get("sbvh_f"))
    u = o.get("sbvh_u").reshape(-1, 3)
    assert np.array_equal(u[:, 0], sb.scene_nodes["first_index"]) and np.array_equal(u[:, 1], sb.scene_nodes["leaf_count"])
    assert np.array_equal(u[:, 2], sb.scene_nodes["depth"])
    assert np.array_equal(sb.leaf_ids, o.get("sbvh_leaf"))
    for k, m in enumerate(sb.meshes):
        assert bits_equal(m.triangles.view(np.float32).reshape(-1), o.get(f"mesh{k}_tri"))
        assert bits_equal(np.concatenate([m.nodes["min"], m.nodes["max"]], 1).reshape(-1), o.get(f"mesh{k}_bvh_f"))
        assert bits_equal(m.aabb, o.get(f"mesh{k}_aabb"))
    p = sb.params
    proj = np.array(list(p.inverted_camera_rotation) + list(p.start) + list(p.right) + list(p.down) +
                    list(p.camera_position) + list(p.C_start) + [p.squared_distance_to_projection_plane, p.sample_size],
                    np.float32)
    assert bits_equal(proj, o.get("proj"))


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("mode", [scenes.MODE_BEAUTY, scenes.MODE_NORMALS, scenes.MODE_NORMAL_MAP, scenes.MODE_DEPTH,
                                  scenes.MODE_MIP_LEVEL, scenes.MODE_UVS])
def test_cpu_context_equals_oracle(name, mode):
    spec = SCENES[name]()
    r = runtime.Renderer(spec, "cpu")
    o = oracle.OracleScene(spec)
    px, dep = r.render(mode=mode)
    col, odep, cnt = o.render(mode=mode)
    assert bits_equal(px[..., :3], col) and bits_equal(dep, odep)
    assert np.all(px[..., 3] == 1.0)
    assert r.count_work(mode=mode) == cnt


def test_cpu_context_ray_cones_mode_equals_oracle():
    spec = SCENES["mesh"]()
    r = runtime.Renderer(spec, "cpu")
    o = oracle.OracleScene(spec, ray_cones=True)
    for mode in (scenes.MODE_BEAUTY, scenes.MODE_MIP_LEVEL):
        px, _ = r.render(mode=mode, ray_cones=True)
        col, _, _ = o.render(mode=mode)
        assert bits_equal(px[..., :3], col)
    # with cones the mip level varies over the frame; in the reference data flow it is always 0
    px, _ = r.render(mode=scenes.MODE_MIP_LEVEL, ray_cones=True)
    assert len(np.unique(px[..., :3].reshape(-1, 3), axis=0)) > 2


def test_ssaa_quad_layout():
    spec = scenes.implicit_scene(0, 48, 27)
    spec.ssaa = True
    r = runtime.Renderer(spec, "cpu")
    o = oracle.OracleScene(spec)
    px, dep = r.render()
    col, odep, _ = o.render()   # (2H, 2W) samples, row-major
    H, W = 27, 48
    quads = col.reshape(H, 2, W, 2, 3).transpose(0, 2, 1, 3, 4).reshape(-1, 3)   # canvas.h:147 pixel quads
    assert bits_equal(px[:, :3], quads)


# --- error paths (CPU context: no GPU needed) ----------------------------------------------------------------
def test_render_before_upload_is_state_error():
    ctx = ctypes.c_void_p()
    assert runtime.lib().slimrt_create(ctypes.byref(ctx), runtime.SLIMRT_DEVICE_CPU) == 0
    p = runtime.RenderParams()
    p.width, p.height, p.max_depth, p.render_mode = 8, 8, 1, 2
    buf = np.zeros(64 * 4, np.float32)
    assert runtime.lib().slimrt_render(ctx, ctypes.byref(p), runtime._ptr(buf), runtime._ptr(buf)) == -4
    assert b"no scene" in runtime.lib().slimrt_last_error(ctx)
    runtime.lib().slimrt_destroy(ctx)


def test_cpu_context_has_no_tile_costs():
    """The CPU context renders rows on host threads: no tiles, no costs; the tile-order flag changes nothing."""
    spec = scenes.implicit_scene(0, 48, 27)
    r = runtime.Renderer(spec, "cpu")
    p = r.params()
    a = r.render(p)
    p.flags = runtime.TILE_ORDER_FIXED
    b = r.render(p)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    costs, order = r.tile_costs()
    assert costs.size == 0 and order.size == 0
    n, res = ctypes.c_uint32(7), ctypes.c_uint32(7)
    assert runtime.lib().slimrt_tile_costs(r.ctx, None, None, 0, ctypes.byref(n), ctypes.byref(res)) == 0
    assert n.value == 0 and res.value == 0
    assert runtime.lib().slimrt_tile_costs(r.ctx, None, None, 0, None, None) == -1   # SLIMRT_E_ARG


def test_invalid_uploads_are_rejected():
    spec = scenes.implicit_scene(0, 32, 18)
    sb = runtime.build_scene(spec)
    r = runtime.Renderer(sb, "cpu")
    bad = sb.scene_nodes.copy()
    bad[0]["first_index"] = 0                        # an inner node pointing at itself
    rc = runtime.lib().slimrt_upload_scene_bvh(r.ctx, runtime._ptr(bad), len(bad), runtime._ptr(sb.leaf_ids), sb.scene_height)
    assert rc == -1
    leaf = sb.leaf_ids.copy()
    leaf[0] = 99
    rc = runtime.lib().slimrt_upload_scene_bvh(r.ctx, runtime._ptr(sb.scene_nodes), len(sb.scene_nodes), runtime._ptr(leaf), sb.scene_height)
    assert rc == -1
    p = r.params()
    p.skybox_color = 3                                # no textures in this scene
    with pytest.raises(runtime.SlimrtError) as e:
        r.render(p)
    assert e.value.code == -1
    geos = sb.geometries.copy()[:-1]
    assert runtime.lib().slimrt_upload_geometries(r.ctx, runtime._ptr(geos), len(geos)) == -1


def test_stack_overflow_is_reported():
    spec = scenes.mesh_pbr_scene(0, 32, 18, level=3, texture_size=16, sky_size=8, irradiance_size=4)
    sb = runtime.build_scene(spec)
    sb.meshes[0].height = 1                           # lie about the mesh BVH height: stacks too small
    sb.mesh_descs[0].height = 1
    r = runtime.Renderer(sb, "cpu")
    with pytest.raises(runtime.SlimrtError) as e:
        r.render()
    assert e.value.code == -5


def test_gpu_context_requires_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    ctx = ctypes.c_void_p()
    assert runtime.lib().slimrt_create(ctypes.byref(ctx), runtime.SLIMRT_DEVICE_GPU) == -2
