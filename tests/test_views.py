"""Device-side scene updates and many views in one launch (include/slimrt_views.h).

GPU checks:
  - slimrt_update_scene_device(REBUILD) leaves on the device exactly the nodes, leaf order and height that the
    host builder (itself bit-exact with the reference, tests/test_oracle_golden.py) derives from the same
    transforms, and the frame equals a frame rendered after a host rebuild and upload, bit for bit;
  - REFIT keeps the uploaded topology, every node box contains its children / primitives, and the frame equals
    the host-rebuilt one (a refitted tree is a valid BVH over the same boxes: the same closest hits);
  - every view of a batch, in every launch form the autotuner can choose (node width 2 or 4 included), is
    bit-identical to its own single-view render with the same node width; ragged view sizes included;
  - graph replays with new transforms each frame equal eager frames, with eager work between replays.
CPU checks: the seeded generators are deterministic and within their stated ranges.
"""
import numpy as np
import pytest

from slimtracincpp_amd import scenes, runtime
from slimtracincpp_amd.scenes import PCG32


def _scene():
    return scenes.mesh_pbr_scene(0, 96, 64, level=3, texture_size=64, sky_size=16, irradiance_size=8)


def test_generators_deterministic_and_in_range():
    spec = _scene()
    a = scenes.instance_transforms(spec, PCG32(5))
    b = scenes.instance_transforms(spec, PCG32(5))
    assert np.array_equal(a, b)
    meshes = [i for i in range(len(a)) if spec.geo_u[i][0] == scenes.GEO_MESH]
    assert len(meshes) == 3
    assert np.array_equal(a[0], spec.geo_f[0])                      # the floor keeps its transform
    for i in meshes:
        s = a[i, 6]
        assert 0.7 <= s <= 1.5 and np.all(a[i, 6:9] == s)
        assert -3.5 <= a[i, 3] <= 3.5 and -3.5 <= a[i, 5] <= 3.5 and np.isclose(a[i, 4], 1.15 * s)
    v = scenes.batched_views(48, 1)
    assert all((x[0], x[1]) == (y[0], y[1]) and np.array_equal(x[3], y[3]) and np.array_equal(x[2], y[2])
               for x, y in zip(v, scenes.batched_views(48, 1)))
    for w, h, eul, pos, focal in v:
        assert (w, h) in scenes.VIEW_RESOLUTIONS
        r = np.linalg.norm(pos - np.array([0, 0.5, 0]))
        assert 6.0 - 1e-4 <= r <= 12.0 + 1e-4
        fov = 2 * np.degrees(np.arctan(1.0 / focal))
        assert 40.0 - 1e-3 <= fov <= 90.0 + 1e-3


# ---------------------------------------------------------------- GPU
def _tensors(p):
    import torch
    n = p.width * p.height * (4 if p.ssaa else 1)
    return (torch.zeros((n, 4), dtype=torch.float32, device="cuda"),
            torch.zeros(n, dtype=torch.float32, device="cuda"))


def _render(r, p):
    import torch
    px, dp = _tensors(p)
    r.render_device(p, px.data_ptr(), dp.data_ptr())
    r.synchronize()
    torch.cuda.synchronize()
    return px.cpu().numpy(), dp.cpu().numpy()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.gpu
def test_device_rebuild_matches_host_build(gpu_available):
    spec = _scene()
    r = runtime.Renderer(spec, "gpu")
    host = runtime.Renderer(spec, "gpu")
    rng = PCG32(11)
    for frame in range(4):
        geo_f = scenes.instance_transforms(spec, rng)
        r.update_scene_device(geo_f, runtime.SCENE_REBUILD)
        host.update_geometries(geo_f)                       # host AABBs + SAH build, then upload
        nodes, leaf = r.download_scene_bvh()
        sc = host.scene
        assert np.array_equal(leaf, sc.leaf_ids), frame
        assert len(nodes) == len(sc.scene_nodes)
        for f in ("min", "max"):
            assert np.array_equal(_bits(nodes[f]), _bits(sc.scene_nodes[f])), (frame, f)
        for f in ("first_index", "leaf_count", "depth"):
            assert np.array_equal(nodes[f], sc.scene_nodes[f]), (frame, f)
        p = r.params()
        a, da = _render(r, p)
        b, db = _render(host, host.params())
        assert np.array_equal(_bits(a), _bits(b)) and np.array_equal(_bits(da), _bits(db)), frame
        assert r.count_work(p) == host.count_work(host.params())


@pytest.mark.gpu
def test_device_refit_is_a_valid_bvh_with_the_same_frame(gpu_available):
    spec = _scene()
    r = runtime.Renderer(spec, "gpu")
    host = runtime.Renderer(spec, "gpu")
    topo, topo_leaf = r.download_scene_bvh()
    rng = PCG32(12)
    for frame in range(4):
        geo_f = scenes.instance_transforms(spec, rng)
        r.update_scene_device(geo_f, runtime.SCENE_REFIT)
        host.update_geometries(geo_f)
        nodes, leaf = r.download_scene_bvh()
        assert np.array_equal(leaf, topo_leaf)
        assert np.array_equal(nodes["first_index"], topo["first_index"])
        assert np.array_equal(nodes["leaf_count"], topo["leaf_count"])
        boxes = host.scene.aabbs                               # host world AABBs of the same transforms
        for i, nd in enumerate(nodes):
            if nd["leaf_count"]:
                kids = [boxes[leaf[nd["first_index"] + k]] for k in range(nd["leaf_count"])]
                lo = np.min([k[:3] for k in kids], 0)
                hi = np.max([k[3:] for k in kids], 0)
            else:
                c0, c1 = nodes[nd["first_index"]], nodes[nd["first_index"] + 1]
                lo = np.minimum(c0["min"], c1["min"])
                hi = np.maximum(c0["max"], c1["max"])
            assert np.array_equal(nd["min"], lo) and np.array_equal(nd["max"], hi), (frame, i)
        a, da = _render(r, r.params())
        b, db = _render(host, host.params())
        assert np.array_equal(_bits(a), _bits(b)) and np.array_equal(_bits(da), _bits(db)), frame


def _views(spec, r, n=7, seed=3):
    cams = scenes.batched_views(n, seed, resolutions=((64, 64), (96, 54), (33, 17), (128, 72), (40, 120)))
    params, bufs = [], []
    for w, h, eul, pos, focal in cams:
        p = r.params(camera=(eul, pos, focal), size=(w, h))
        params.append(p)
        bufs.append(_tensors(p))
    return params, bufs


@pytest.mark.gpu
def test_views_equal_single_view_renders_in_every_launch_form(gpu_available):
    """Node width 2 against single-view renders of the sibling-pair walk, widths 4 and 8 against single-view
    renders with SLIMRT_BVH4 / SLIMRT_BVH8 (the same per-sample code: bit for bit either way)."""
    import torch
    spec = _scene()
    r = runtime.Renderer(spec, "gpu")
    params, bufs = _views(spec, r)
    refs = {2: [_render(r, p) for p in params]}
    for width, flag in ((4, runtime.BVH4), (8, runtime.BVH8)):
        wide = []
        for p in params:
            q = runtime.RenderParams.from_buffer_copy(p)
            q.flags |= flag
            wide.append(_render(r, q))
        refs[width] = wide
    v = runtime.Views(r, params, [b[0].data_ptr() for b in bufs], [b[1].data_ptr() for b in bufs])
    forms = [(shape, 1, 1, 0, w) for w in (2, 4, 8) for shape in range(4)] + \
            [(shape, bpc, grab, 1, w) for w in (2, 4, 8) for shape in range(4) for bpc in (1, 2) for grab in (1, 4)]
    for form in forms + [(0, 2, 1, 0, 0)]:               # node width 0: as the views' flags ask (2 here)
        for px, dp in bufs:
            px.zero_()
            dp.zero_()
        v.configure(*form)
        v.frame()
        r.synchronize()
        torch.cuda.synchronize()
        for i, ((px, dp), (rp, rd)) in enumerate(zip(bufs, refs[form[4] or 2])):
            assert np.array_equal(_bits(px.cpu().numpy()), _bits(rp)), (form, i)
            assert np.array_equal(_bits(dp.cpu().numpy()), _bits(rd)), (form, i)
    cfg, ms = v.autotune(reps=2)
    assert tuple(cfg.values()) in forms and ms > 0
    v.close()


@pytest.mark.gpu
def test_graph_replays_equal_eager_frames(gpu_available):
    import torch
    spec = _scene()
    r_graph = runtime.Renderer(spec, "gpu")
    r_eager = runtime.Renderer(spec, "gpu")
    params, bufs_g = _views(spec, r_graph, n=5, seed=4)
    _, bufs_e = _views(spec, r_eager, n=5, seed=4)
    vg = runtime.Views(r_graph, params, [b[0].data_ptr() for b in bufs_g], [b[1].data_ptr() for b in bufs_g])
    ve = runtime.Views(r_eager, params, [b[0].data_ptr() for b in bufs_e], [b[1].data_ptr() for b in bufs_e])
    rng = PCG32(21)
    for frame in range(6):   # three replays of a refit frame's graph, then three of a rebuild frame's
        geo_f = scenes.instance_transforms(spec, rng)
        mode = runtime.SCENE_REFIT if frame < 3 else runtime.SCENE_REBUILD
        vg.frame(geo_f, mode, graph=True)
        ve.frame(geo_f, mode, graph=False)
        r_graph.synchronize()
        r_eager.synchronize()
        torch.cuda.synchronize()
        scratch = torch.ones(1 << 20, device="cuda") * frame   # eager work between replays
        torch.cuda.synchronize()
        del scratch
        for i, ((pg, dg), (pe, de)) in enumerate(zip(bufs_g, bufs_e)):
            assert np.array_equal(_bits(pg.cpu().numpy()), _bits(pe.cpu().numpy())), (frame, i)
            assert np.array_equal(_bits(dg.cpu().numpy()), _bits(de.cpu().numpy())), (frame, i)
        assert vg.last_ms() > 0
    vg.close()
    ve.close()
