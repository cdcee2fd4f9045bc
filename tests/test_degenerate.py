"""Degenerate inputs the reference meets without complaint: ray directions with exact zero components (Ray::reset
takes 1/0 = inf and 0 * inf = NaN in the slab test's scaled origin, src/slim/core/ray.h:40-49) and geometries with
a zero scale component (Ray::localize divides by it, ray.h:29-38).  The oracle matches the reference's own output
bit for bit (live, through the compiled driver); the CPU context matches the oracle bit for bit; the GPU matches
hit for hit and within the beauty tolerance."""
import os

import numpy as np
import pytest

import oracle
from slimtracincpp_amd import runtime, scenes
from tests.helpers import bits_equal

needs_ref = pytest.mark.skipif(not (os.path.exists(oracle.REF_PATH) and os.path.isdir("/root/reference/src")),
                               reason="compiled reference not available")


def _axis_aligned():
    """An odd-sized canvas seen by an unrotated camera at x = 0: the middle column's rays have direction.x = 0
    from origin.x = 0, the middle row's direction.y = 0 (checked below)."""
    spec = scenes.implicit_scene(0, 63, 35, max_depth=3, transparent=True)
    spec.camera_euler = np.zeros(3, np.float32)
    return spec


def _zero_scale():
    spec = scenes.implicit_scene(2, 48, 32, max_depth=3)
    spec.geo_f = spec.geo_f.copy()
    spec.geo_f[1, 6] = 0.0                 # first non-floor geometry: zero scale along x
    spec.geo_f[2, 8] = 0.0                 # next one: zero scale along z
    return spec


CASES = {"axis_aligned": _axis_aligned, "zero_scale": _zero_scale}


def test_axis_aligned_case_has_zero_components():
    r = runtime.Renderer(_axis_aligned(), "cpu")
    p = r.params()
    start, right, down = (np.array(v[:], np.float32) for v in (p.start, p.right, p.down))
    assert p.camera_position[0] == 0.0
    assert (start + right * np.float32(31))[0] == 0.0 and (start + down * np.float32(17))[1] == 0.0


@needs_ref
@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_equals_reference(name):
    spec = CASES[name]()
    ref = oracle.run_reference(spec)
    col, dep, _ = oracle.OracleScene(spec).render()
    assert bits_equal(col.reshape(-1), ref["color"]) and bits_equal(dep.reshape(-1), ref["depth"])
    geo, tri, _ = oracle.OracleScene(spec).primary_hits()
    assert np.array_equal(geo.reshape(-1), ref["hit_geo"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_cpu_context_equals_oracle(name):
    spec = CASES[name]()
    px, dp = runtime.Renderer(spec, "cpu").render()
    col, dep, _ = oracle.OracleScene(spec).render()
    assert bits_equal(px[..., :3], col) and bits_equal(dp, dep)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("flags", [0, runtime.WAVEFRONT | runtime.WAVEFRONT_SPLIT | runtime.BVH4])
def test_gpu_equals_oracle(gpu_available, name, flags):
    spec = CASES[name]()
    r = runtime.Renderer(spec, "gpu")
    p = r.params()
    p.flags = flags
    px, dp = r.render(p)
    o = oracle.OracleScene(spec)
    col, dep, _ = o.render()
    fin = np.isfinite(col)
    assert np.array_equal(fin, np.isfinite(px[..., :3]))
    assert np.max(np.abs(px[..., :3][fin] - col[fin]), initial=0.0) <= 1e-4
    geo, _, _ = o.primary_hits()
    g, _, _ = r.primary_hits()
    assert np.array_equal(g, geo)
