"""The SAH builder's quicksort (BVHPartition::partition, src/slim/scene/bvh_builder.h:15-78: iterative Lomuto,
pivot the last element, strict <) on ranges whose keys never decrease: its result is every maximal run of equal
keys rotated right by one.  The host builder and the oracle compute that result in one pass instead of the
procedure's quadratic steps.  Pinned here three ways: the procedure itself (step by step, in Python) against the
formula on random tie-heavy sorted keys; the oracle's builder with and without the one-pass form on adversarial
boxes and on meshes; and the product's host builder against the oracle's step-by-step builder."""
import numpy as np
import pytest

import oracle
from slimtracincpp_amd import runtime, scenes
from tests.helpers import bits_equal


def lomuto(keys, ids):
    """BVHPartition::partition's sort, step by step (bvh_builder.h:22-78)."""
    ids = list(ids)
    stack = [0, len(ids) - 1]
    while stack:
        end, start = stack.pop(), stack.pop()
        p = start
        pivot = keys[ids[end]]
        for i in range(start, end):
            if keys[ids[i]] < pivot:
                ids[i], ids[p] = ids[p], ids[i]
                p += 1
        ids[end], ids[p] = ids[p], ids[end]
        if p - 1 > start:
            stack += [start, p - 1]
        if p + 1 < end:
            stack += [p + 1, end]
    return ids


def rotated_runs(keys, ids):
    out, i = [], 0
    while i < len(ids):
        j = i
        while j + 1 < len(ids) and not keys[ids[i]] < keys[ids[j + 1]]:
            j += 1
        out += [ids[j]] + ids[i:j]
        i = j + 1
    return out


@pytest.mark.parametrize("seed", range(40))
def test_sorted_range_result_is_rotated_runs(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 120))
    keys = np.sort(rng.integers(0, max(2, n // int(rng.integers(1, 6))), n)).astype(np.float32)
    if seed % 5 == 0:
        keys[:] = 1.0                                   # one run
    if seed % 7 == 0:
        keys[rng.integers(0, n, n // 3)] = -0.0         # signed zeros compare equal
        keys = np.sort(keys)
    ids = list(rng.permutation(1000)[:n])               # arbitrary ids: only positions matter
    k = {i: float(v) for i, v in zip(ids, keys)}
    assert lomuto(k, ids) == rotated_runs(k, ids)


def _tie_heavy_boxes(seed, n, sorted_axis):
    rng = np.random.default_rng(seed)
    lo = rng.integers(0, 40, (n, 3)).astype(np.float32) * np.float32(0.25)
    hi = lo + rng.integers(0, 4, (n, 3)).astype(np.float32) * np.float32(0.25)
    if sorted_axis is not None:
        order = np.argsort(hi[:, sorted_axis], kind="stable")
        lo, hi = lo[order], hi[order]
    return np.concatenate([lo, hi], 1)


@pytest.mark.parametrize("seed,n,sorted_axis", [(0, 300, 1), (1, 2000, 0), (2, 5000, None), (3, 700, 2), (4, 64, 1)])
def test_oracle_builder_one_pass_equals_step_by_step(seed, n, sorted_axis):
    boxes = _tie_heavy_boxes(seed, n, sorted_axis)
    for leaf in (1, 2, 4):
        oracle.set_plain_quicksort(True)
        try:
            plain = oracle.build_bvh(boxes, leaf)
        finally:
            oracle.set_plain_quicksort(False)
        fast = oracle.build_bvh(boxes, leaf)
        assert all(bits_equal(a, b) for a, b in zip(plain[:3], fast[:3])) and plain[3] == fast[3]
        # the product's host builder (one-pass form as well) against the step-by-step procedure
        nodes, leaf_ids, height = runtime.build_bvh(boxes, leaf)
        f = np.concatenate([nodes["min"], nodes["max"]], 1)
        assert bits_equal(f, plain[0]) and np.array_equal(leaf_ids, plain[2]) and height == plain[3]
        assert np.array_equal(nodes["first_index"], plain[1][:, 0]) and np.array_equal(nodes["leaf_count"], plain[1][:, 1])


@pytest.mark.parametrize("make", [
    lambda: scenes.large_scene(1, 16, 16, uv_sphere=(33, 64), texture_size=16, sky_size=8, irradiance_size=4),
    lambda: scenes.mesh_pbr_scene(2, 16, 16, level=4, texture_size=16, sky_size=8, irradiance_size=4, leaf_size=2),
], ids=["uv_sphere_leaf4", "icosphere_leaf2"])
def test_oracle_mesh_build_one_pass_equals_step_by_step(make):
    spec = make()
    oracle.set_plain_quicksort(True)
    try:
        plain = oracle.OracleScene(spec)
    finally:
        oracle.set_plain_quicksort(False)
    fast = oracle.OracleScene(spec)
    for k in ("mesh0_bvh_f", "mesh0_bvh_u", "mesh0_tri", "sbvh_f", "sbvh_u", "sbvh_leaf"):
        assert bits_equal(plain.get(k), fast.get(k)), k
