"""8-wide mesh BVHs with quantised child boxes (render flag SLIMRT_BVH8; slimrt_host_collapse_bvh8 builds the same
trees the upload does).

The collapse, against the sibling-pair BVH (the reference's shipped monkey.mesh and cube.mesh rebuilt bit for bit,
and generated meshes with leaves of 1 and 4 triangles): every node is reached once from the root; its leaf records
are exactly the BVH2's leaves (exact box, triangle range, the triangles' test records), so they partition the
triangles; every decoded child box contains the exact child box with at least two ulps to spare on every bound
(what makes the walk's fused slab test accept every ray the reference's test on the exact box accepts); the stack
bound covers a walk that pushes every inner child.  Frames: the CPU context (same per-sample code) against the
oracle bit for bit, the GPU against the sibling-pair walk and the oracle within the parity bars, in both tile orders,
at the benchmark's full size, with the costliest tiles as quarter and half waves."""
import os

import numpy as np
import pytest

import oracle
from slimtracincpp_amd import assets, runtime, scenes
from tests.helpers import GOLDEN, bits_equal, mesh_spec_from_asset


def _meshes():
    out = {}
    for fname, leaf in (("monkey.mesh", 2), ("cube.mesh", 4)):
        m = assets.read_mesh(os.path.join(GOLDEN, fname))
        out[fname] = mesh_spec_from_asset(m, leaf).meshes[0]
    out["icosphere_l5_leaf4"] = scenes.mesh_pbr_scene(0, 16, 16, level=5, texture_size=8, sky_size=4,
                                                      irradiance_size=4).meshes[0]
    out["icosphere_l3_leaf1"] = scenes.mesh_pbr_scene(2, 16, 16, level=3, leaf_size=1, texture_size=8, sky_size=4,
                                                      irradiance_size=4).meshes[0]
    return out


MESHES = _meshes()


def _decode(node, slot):
    """(lo xyz, hi xyz) of a slot's decoded box: origin + q * step per axis, rounded once to float32 (the walk's
    fused multiply-add; q * step is exact, so the float64 sum rounded to float32 is the same number)."""
    steps = [np.float64(2.0) ** (int((int(node["exps"]) >> (8 * a)) & 0xFF) - 127) for a in range(3)]
    lo = [np.float32(np.float64(node["origin"][a]) + np.float64(node["qlo"][a][slot]) * steps[a]) for a in range(3)]
    hi = [np.float32(np.float64(node["origin"][a]) + np.float64(node["qhi"][a][slot]) * steps[a]) for a in range(3)]
    return np.array(lo, np.float32), np.array(hi, np.float32)


def _ulp(x):
    x = np.abs(np.float32(x))
    return np.float64(np.nextafter(x, np.float32(np.inf))) - np.float64(x)


def _collapse(name):
    mb = runtime.MeshBuild(MESHES[name])
    nodes8, leaves, stack = runtime.collapse_bvh8(mb.nodes, mb.triangles)
    return mb, nodes8, leaves.view(np.uint32).reshape(-1, 4), stack


@pytest.mark.parametrize("name", sorted(MESHES))
def test_collapse8_keeps_leaves_and_contains_boxes(name):
    mb, n8, units, stack = _collapse(name)
    n2 = mb.nodes
    leaf2 = {(int(n["first_index"]), int(n["leaf_count"])): np.concatenate([n["min"], n["max"]]).astype(np.float32)
             for n in n2 if n["leaf_count"]}
    covered = np.zeros(len(mb.triangles), np.int32)
    seen = np.zeros(len(n8), np.int32)
    tri_rec = np.zeros((len(mb.triangles), 12), np.float32)
    lt = mb.triangles["local_to_tangent"].reshape(-1, 3, 3)
    tri_rec[:, 0], tri_rec[:, 1], tri_rec[:, 2], tri_rec[:, 3] = lt[:, 0, 0], lt[:, 0, 1], lt[:, 1, 0], lt[:, 1, 1]
    tri_rec[:, 4], tri_rec[:, 5] = lt[:, 2, 0], lt[:, 2, 1]
    tri_rec[:, 6:9], tri_rec[:, 9:12] = mb.triangles["position"], mb.triangles["normal"]

    def exact_box(k):   # a slot's exact box: the leaf record's, or the union of everything below an inner slot
        return below[k]

    below = {}

    def visit(i):
        seen[i] += 1
        node = n8[i]
        lo_all, hi_all = np.full(3, np.inf, np.float32), np.full(3, -np.inf, np.float32)
        for s in range(8):
            cnt = int(node["count"][s])
            if cnt == 255:
                continue
            if cnt:
                u = int(node["child"][s])
                head = units[u].view(np.float32)
                first, count = int(units[u][3]), int(units[u + 1][3])
                box = np.concatenate([head[:3], units[u + 1].view(np.float32)[:3]])
                assert count == cnt and (first, count) in leaf2, (name, i, s)
                assert bits_equal(box, leaf2[(first, count)])
                recs = units[u + 2:u + 2 + 3 * count].view(np.float32).reshape(count, 12)
                assert bits_equal(recs, tri_rec[first:first + count])
                covered[first:first + count] += 1
            else:
                box = visit(int(node["child"][s]))
            lo, hi = _decode(node, s)
            for a in range(3):   # two ulps of the exact bound to spare
                assert np.float64(lo[a]) <= np.float64(box[a]) - 2 * _ulp(box[a]), (name, i, s, a)
                assert np.float64(hi[a]) >= np.float64(box[3 + a]) + 2 * _ulp(box[3 + a]), (name, i, s, a)
            lo_all, hi_all = np.minimum(lo_all, box[:3]), np.maximum(hi_all, box[3:])
        return np.concatenate([lo_all, hi_all])

    visit(0)
    assert np.all(seen == 1)              # a tree: every node reached exactly once
    assert np.all(covered == 1)           # the leaves partition the triangles
    assert len(units) >= 6 and not np.any(units[-6:])   # a record's worth of padding at the end


@pytest.mark.parametrize("name", sorted(MESHES))
def test_stack_bound8_covers_pushing_every_inner_child(name):
    mb, n8, units, stack = _collapse(name)
    worst = 0
    st = [0]
    while st:   # visit = pop, push every inner child: the deepest the walk's stack can get
        node = n8[st.pop()]
        for s in range(8):
            if int(node["count"][s]) == 0:
                st.append(int(node["child"][s]))
        worst = max(worst, len(st))
    assert worst <= stack


def test_collapse8_of_a_single_leaf():
    m = scenes.MeshSpec(positions=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32),
                        position_indices=np.array([[0, 1, 2]], np.uint32), leaf_size=4)
    mb = runtime.MeshBuild(m)
    n8, leaves, stack = runtime.collapse_bvh8(mb.nodes, mb.triangles)
    assert len(n8) == 1 and stack == 0
    assert sorted(int(c) for c in n8[0]["count"]) == [1] + [255] * 7


SCENES = {
    "mesh_pbr_l5": lambda: scenes.mesh_pbr_scene(0, 160, 90, level=5, texture_size=64, sky_size=16, irradiance_size=8),
    "mesh_leaf1_d8": lambda: scenes.mesh_pbr_scene(2, 128, 72, max_depth=8, level=4, leaf_size=1, texture_size=64,
                                                   sky_size=16, irradiance_size=8),
    "lit_mesh": lambda: scenes.lit_mesh_scene(0, 128, 72, level=3, texture_size=32, sky_size=8, irradiance_size=4),
    "large_small": lambda: scenes.large_scene(0, 128, 72, uv_sphere=(33, 64), texture_size=64, sky_size=16,
                                              irradiance_size=8),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_cpu_context_bvh8_equals_oracle(name):
    """The 8-wide walk in the CPU context (the kernels' per-sample code): the oracle's frame bit for bit (the lit
    scene's shadow rays take the walk's any-hit exit), fewer node visits than the 4-wide walk."""
    spec = SCENES[name]()
    col, dep, cnt = oracle.OracleScene(spec).render()
    r = runtime.Renderer(spec, "cpu")
    p = r.params()
    p.flags = runtime.BVH8
    px, dp = r.render(p)
    assert bits_equal(px[..., :3], col) and bits_equal(dp, dep)
    c8 = r.count_work(p)
    p.flags = runtime.BVH4
    c4 = r.count_work(p)
    r.close()
    assert c8["rays"] == cnt["rays"] and c8["mesh_hits"] == cnt["mesh_hits"]
    assert c8["pairs"] < c4["pairs"]


def test_bvh8_leaf_limit():
    """Leaves of more than 254 triangles do not fit an 8-wide node's count byte: such meshes render with node
    width 2 or 4 and SLIMRT_BVH8 frames are refused."""
    spec = scenes.mesh_pbr_scene(0, 32, 18, level=2, leaf_size=320, texture_size=8, sky_size=4, irradiance_size=4)
    r = runtime.Renderer(spec, "cpu")
    p = r.params()
    p.flags = runtime.BVH8
    with pytest.raises(runtime.SlimrtError, match="254"):
        r.render(p)
    p.flags = runtime.BVH4
    r.render(p)
    r.close()


# ---------------------------------------------------------------------------------------------------- GPU
def _frame(r, p):
    import torch
    n = p.width * p.height * (4 if p.ssaa else 1)
    px = torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda")
    dp = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    r.render_device(p, px.data_ptr(), dp.data_ptr())
    r.synchronize()
    torch.cuda.synchronize()
    return px.cpu().numpy(), dp.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SCENES))
@pytest.mark.parametrize("order", [0, runtime.TILE_ORDER_FIXED])
def test_gpu_bvh8_equals_oracle(gpu_available, name, order):
    spec = SCENES[name]()
    col, dep, cnt = oracle.OracleScene(spec).render()
    r = runtime.Renderer(spec, "gpu")
    p = r.params(output=runtime.OUTPUT_RADIANCE)
    p.flags = runtime.BVH8 | order
    for _ in range(4):   # from the fourth frame in cost order
        px, dp = _frame(r, p)
    got = px.reshape(spec.height, spec.width, 4)[..., :3]
    err = np.abs(got - col)
    print(name, order, "max err", float(err.max()), "bit-exact", float(np.mean(got.view(np.uint32) == col.view(np.uint32))))
    assert np.mean(np.all(err <= 1e-4, axis=-1)) >= 0.9999
    assert np.array_equal(np.isfinite(dp), np.isfinite(dep.reshape(-1)))
    assert r.count_work(p)["rays"] == cnt["rays"]
    cpu = runtime.Renderer(spec, "cpu")
    assert cpu.count_work(p) == r.count_work(p)   # the same walk on both sides: the same work
    cpu.close()
    r.close()


@pytest.mark.gpu
def test_gpu_bvh8_benchmark_size_against_oracle(gpu_available):
    """The 1080p mesh + PBR frame with 8-wide nodes, tiles costliest first from the fourth frame (the costliest 31
    as four quarter waves, the next as two half waves), frames 1-4 against the oracle: beauty within 1e-4 on every
    sample, depth within 1e-5 relative, primary hit ids on >= 99.99% of samples; every tile written (NaN-filled
    canvases)."""
    spec = scenes.mesh_pbr_scene(0, 1920, 1080)
    o = oracle.OracleScene(spec)
    col, dep, _ = o.render()
    r = runtime.Renderer(spec, "gpu")
    p = r.params(output=runtime.OUTPUT_RADIANCE)
    p.flags = runtime.BVH8
    for frame in range(5):
        px, dp = _frame(r, p)
        got = px.reshape(spec.height, spec.width, 4)[..., :3]
        err = float(np.max(np.abs(got - col)))
        print(f"frame {frame}: max err {err:.3g}, bit-exact {np.mean(got.view(np.uint32) == col.view(np.uint32)):.6f}")
        assert err <= 1e-4, (frame, err)
        d, d2 = dp.reshape(spec.height, spec.width), dep
        assert np.array_equal(np.isfinite(d), np.isfinite(d2))
        fin = np.isfinite(d2)
        assert np.all(np.abs(d[fin] - d2[fin]) <= 1e-5 * np.abs(d2[fin]))
    costs, order = r.tile_costs()
    assert not np.array_equal(order, np.arange(order.size))   # the last frames ran their costliest tiles first
    p.output = runtime.OUTPUT_HITS
    hits, _ = r.render(p)
    g, t = hits[..., 0].view(np.int32), hits[..., 1].view(np.int32)
    g2, t2, _ = o.primary_hits()
    agree = (g == g2) & (t == t2)
    print(f"8-wide 1080p hit-id agreement {agree.mean():.6f}, mismatches {int((~agree).sum())}")
    assert agree.mean() >= 0.9999
    r.close()


@pytest.mark.gpu
def test_gpu_bvh8_refused_by_other_forms(gpu_available):
    spec = SCENES["mesh_pbr_l5"]()
    r = runtime.Renderer(spec, "gpu")
    p = r.params()
    p.flags = runtime.BVH8 | runtime.WAVEFRONT
    with pytest.raises(runtime.SlimrtError, match="wavefront"):
        r.render(p)
    r.close()
