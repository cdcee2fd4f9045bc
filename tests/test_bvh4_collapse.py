"""The 4-wide collapse of mesh BVHs (slimrt_host_collapse_bvh4: the trees SLIMRT_BVH4 frames walk, built by the
same host code at upload), on the reference's shipped monkey.mesh (leaf 2) and cube.mesh (leaf 4) rebuilt bit for
bit, and on generated meshes with leaves of 1 and 4 triangles.  Invariants checked against the sibling-pair BVH:
every 4-wide node is reached once from the root; its used slots come first, two to four of them, fewer than four
only when all are leaves; an inner slot carries the box of a BVH2 inner node and contains every box below it; the
leaf slots are exactly the BVH2's leaves (same triangle ranges, same boxes), so they partition the triangles; the
height is the longest root-to-leaf chain and no more than the BVH2's."""
import os

import numpy as np
import pytest

from slimtracincpp_amd import assets, runtime, scenes
from tests.helpers import GOLDEN, mesh_spec_from_asset


def _meshes():
    out = {}
    for fname, leaf in (("monkey.mesh", 2), ("cube.mesh", 4)):
        m = assets.read_mesh(os.path.join(GOLDEN, fname))
        out[fname] = mesh_spec_from_asset(m, leaf).meshes[0]
    out["icosphere_l4_leaf4"] = scenes.mesh_pbr_scene(0, 16, 16, level=4, texture_size=8, sky_size=4,
                                                      irradiance_size=4).meshes[0]
    out["icosphere_l3_leaf1"] = scenes.mesh_pbr_scene(2, 16, 16, level=3, leaf_size=1, texture_size=8, sky_size=4,
                                                      irradiance_size=4).meshes[0]
    return out


MESHES = _meshes()


def _box(n):
    return np.concatenate([n["min"], n["max"]]).astype(np.float32)


def _slot_box(n4, j):
    return np.array([n4["min"][0][j], n4["min"][1][j], n4["min"][2][j],
                     n4["max"][0][j], n4["max"][1][j], n4["max"][2][j]], np.float32)


def _key(box):
    return box.tobytes()


@pytest.mark.parametrize("name", sorted(MESHES))
def test_collapse_keeps_leaves_and_boxes(name):
    b = runtime.MeshBuild(MESHES[name])
    n2, T = b.nodes, len(b.triangles)
    n4, h4 = runtime.collapse_bvh4(n2, T)
    assert 1 <= len(n4) <= (len(n2) + 1) // 2
    leaves2 = {(int(n["first_index"]), int(n["leaf_count"])): _key(_box(n)) for n in n2 if n["leaf_count"]}
    inner2 = {_key(_box(n)) for n in n2 if not n["leaf_count"]}
    seen, leaf_slots = np.zeros(len(n4), int), []

    def walk(i, depth):
        """Returns the box union of everything below node i; checks the node's slots."""
        seen[i] += 1
        nd = n4[i]
        used = [j for j in range(4) if nd["child"][j] != 0xFFFFFFFF]
        assert used == list(range(len(used))), i                      # used slots first
        assert 2 <= len(used) <= 4 or (i == 0 and len(n2) == 1 and len(used) == 1), i
        for j in range(len(used), 4):                                   # unused: empty box, no range
            assert nd["count"][j] == 0
            assert np.all(_slot_box(nd, j)[:3] == np.inf) and np.all(_slot_box(nd, j)[3:] == -np.inf)
        if len(used) < 4:
            assert all(nd["count"][j] > 0 for j in used), i               # only leaves stop the expansion
        height = depth
        union = np.array([np.inf] * 3 + [-np.inf] * 3, np.float32)
        for j in used:
            box = _slot_box(nd, j)
            if nd["count"][j]:
                rng = (int(nd["child"][j]), int(nd["count"][j]))
                assert leaves2.get(rng) == _key(box), (i, j)
                leaf_slots.append(rng)
            else:
                assert _key(box) in inner2, (i, j)
                c = int(nd["child"][j])
                assert i < c < len(n4), (i, j)
                below, h = walk(c, depth + 1)
                assert np.all(box[:3] <= below[:3]) and np.all(box[3:] >= below[3:]), (i, j)
                height = max(height, h)
            union = np.concatenate([np.minimum(union[:3], box[:3]), np.maximum(union[3:], box[3:])])
        return union, height

    _, h = walk(0, 1)
    assert np.all(seen == 1)                                            # a tree: every node reached once
    assert h == h4 and h4 <= b.height
    assert sorted(leaf_slots) == sorted(leaves2)                       # exactly the BVH2 leaves ...
    starts = sorted(leaf_slots)
    assert starts[0][0] == 0 and all(a[0] + a[1] == c[0] for a, c in zip(starts, starts[1:]))
    assert starts[-1][0] + starts[-1][1] == T                           # ... which partition the triangles


def _all_hit_walk(n4, rng):
    """Deepest stack of the 4-wide walk when every box is hit, visiting each node's inner children in a random
    order (the nearest-first order of some ray); the walk pushes them all and pops one."""
    stack, deepest, cur = [], 0, 0
    while True:
        inner = [int(n4[cur]["child"][j]) for j in range(4)
                 if n4[cur]["child"][j] != 0xFFFFFFFF and n4[cur]["count"][j] == 0]
        rng.shuffle(inner)
        stack.extend(inner)
        deepest = max(deepest, len(stack))
        if not stack:
            return deepest
        cur = stack.pop()


@pytest.mark.parametrize("name", sorted(MESHES))
def test_stack_bound_covers_every_visiting_order(name):
    """The stack capacity the renderer sizes from the collapse holds the all-boxes-hit walk in every visiting order
    tried (60 random orders) and stays within three entries per level (the bound it replaces)."""
    b = runtime.MeshBuild(MESHES[name])
    n4, h4, bound = runtime.collapse_bvh4(b.nodes, len(b.triangles), with_stack=True)
    rng = np.random.default_rng(5)
    walks = [_all_hit_walk(n4, rng) for _ in range(60)]
    assert max(walks) <= bound <= 3 * h4 + 4


def test_collapse_of_a_single_leaf():
    n2 = np.zeros(1, runtime.BVH_NODE)
    n2[0]["min"], n2[0]["max"], n2[0]["first_index"], n2[0]["leaf_count"] = (-1, -1, -1), (1, 1, 1), 0, 3
    n4, h4 = runtime.collapse_bvh4(n2, 3)
    assert len(n4) == 1 and h4 == 1
    assert n4[0]["child"][0] == 0 and n4[0]["count"][0] == 3 and list(n4[0]["child"][1:]) == [0xFFFFFFFF] * 3


def test_collapse_rejects_bad_indices():
    b = runtime.MeshBuild(MESHES["cube.mesh"])
    bad = b.nodes.copy()
    bad[0]["first_index"] = 0                                           # a child before its parent: a cycle
    with pytest.raises(runtime.SlimrtError):
        runtime.collapse_bvh4(bad, len(b.triangles))
    with pytest.raises(runtime.SlimrtError):
        runtime.collapse_bvh4(b.nodes, len(b.triangles) - 1)           # a leaf range past the triangles
