"""The lighting branches of the mesh example (src/examples/07_Meshes.cpp:40-110), on the oracle: the Phong BRDF
(radianceFraction, src/slim/renderer/surface_shader.h:130-142), directional lights (isFacingLight,
surface_shader.h:147-152), shadow rays that end on a mesh (MeshTracer::trace in any-hit mode,
src/slim/scene/mesh_tracer.h:43, 76, 90) and lights that cast no shadows (shadeFromLight, surface_shader.h:51-59).

The oracle's frames of these scenes equal the reference's own (tests/test_oracle_golden.py: live cases
lit_mesh_* and the ref_lit_mesh_* goldens).  The checks here show each branch changes the frame, so those
bit-exact comparisons are comparisons of the branch, not of code that never runs."""
import numpy as np

import oracle
from slimtracincpp_amd import scenes


def _scene(**kw):
    return scenes.lit_mesh_scene(1, 96, 54, level=3, texture_size=32, sky_size=8, irradiance_size=4, **kw)


def _changed(a, b, tol=1e-4):
    return int(np.sum(np.any(np.abs(a - b) > tol, axis=-1)))


def test_scene_uses_every_branch():
    s = _scene()
    assert scenes.BRDF_PHONG in s.mat_u[:, 4] and scenes.BRDF_BLINN in s.mat_u[:, 4]
    assert np.any(s.light_u & scenes.LIGHT_DIRECTIONAL)
    assert np.any((s.light_u & scenes.LIGHT_SHADOWING) == 0)
    mesh = s.geo_u[:, 0] == scenes.GEO_MESH
    assert np.any(mesh & ((s.geo_u[:, 3] & scenes.GEO_SHADOWING) != 0))
    d = s.light_f[s.light_u & scenes.LIGHT_DIRECTIONAL != 0, 3:6]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6) and np.all(d[:, 1] < 0)


def test_mesh_shadow_rays_change_the_frame():
    """Shadow rays from the floor and the implicit primitives end on the Phong and Blinn meshes: clearing the
    meshes' GEOMETRY_IS_SHADOWING flag brightens those samples and cuts the triangle tests."""
    base = _scene(ibl=False)
    col, _, cnt = oracle.OracleScene(base).render()
    lit = _scene(ibl=False)
    lit.geo_u[lit.geo_u[:, 0] == scenes.GEO_MESH, 3] = scenes.GEO_VISIBLE
    col2, _, cnt2 = oracle.OracleScene(lit).render()
    assert _changed(col, col2) > 100
    assert np.sum(col2 - col) > 0                  # only brighter: a shadow was lifted
    assert cnt2["rays"] == cnt["rays"] and cnt2["triangles"] < cnt["triangles"]


def test_directional_light_contributes():
    base = _scene(ibl=False)
    col, _, _ = oracle.OracleScene(base).render()
    off = _scene(ibl=False)
    off.light_f[off.light_u & scenes.LIGHT_DIRECTIONAL != 0, 6] = 0.0
    col2, _, _ = oracle.OracleScene(off).render()
    assert _changed(col, col2) > 1000


def test_directional_light_casts_shadows():
    base = _scene(ibl=False)
    col, _, _ = oracle.OracleScene(base).render()
    free = _scene(ibl=False)
    free.light_u[free.light_u & scenes.LIGHT_DIRECTIONAL != 0] = scenes.LIGHT_DIRECTIONAL
    col2, _, _ = oracle.OracleScene(free).render()
    assert _changed(col, col2) > 100 and np.sum(col2 - col) > 0


def test_phong_differs_from_blinn():
    """The Phong branch (exponent 4 on R.L) is not the Blinn one (exponent 16 on N.H): swapping the Phong
    materials to Blinn changes the samples that see them."""
    base = _scene()
    col, _, _ = oracle.OracleScene(base).render()
    swapped = _scene()
    swapped.mat_u[swapped.mat_u[:, 4] == scenes.BRDF_PHONG, 4] = scenes.BRDF_BLINN
    col2, _, _ = oracle.OracleScene(swapped).render()
    g, _, _ = oracle.OracleScene(base).primary_hits()
    phong_geo = np.isin(g, np.nonzero(base.mat_u[base.geo_u[:, 1], 4] == scenes.BRDF_PHONG)[0])
    assert _changed(col[phong_geo], col2[phong_geo]) > 0.5 * phong_geo.sum()


def test_phong_specular_closed_form():
    """A Phong surface under one unshadowed point light and nothing else: the sample equals
    tone_map(albedo * roughness / pi * c + F0 * max(R.L, 0)^4 * (1 - roughness) * c), c = NdotL * I / Ld^2,
    evaluated in float64 from the hit distance the oracle reports (5e-5: float32 rounding of the terms, steepened
    by the tone map's slope of up to 8.3 near black)."""
    s = scenes.lit_mesh_scene(0, 64, 36, level=2, ibl=False, max_depth=1)
    # one Phong quad floor facing up, one light straight above the camera's target
    s.geo_f = s.geo_f[:1].copy()
    s.geo_u = s.geo_u[:1].copy()
    s.geo_u[0, 1] = 2                                 # the Phong material
    s.light_f = np.array([[1.0, 1.0, 1.0, 0.5, 3.0, 1.0, 12.0]], np.float32)
    s.light_u = np.array([0], np.uint32)
    s.meshes = []
    o = oracle.OracleScene(s)
    col, dep, _ = o.render()
    g, _, t = o.primary_hits()
    from slimtracincpp_amd import runtime
    rot = runtime.camera_rotation(s.camera_euler).astype(np.float64)
    cam = s.camera_position.astype(np.float64)
    m = s.mat_f[2]
    albedo, rough, f0 = m[0:3].astype(np.float64), float(m[3]), m[11:14].astype(np.float64)
    lp, inten = s.light_f[0, 3:6].astype(np.float64), float(s.light_f[0, 6])
    focal = float(s.focal_length)
    checked = 0
    for y in range(0, 36, 5):
        for x in range(0, 64, 7):
            if g[y, x] != 0:
                continue
            # the camera ray through the sample centre (CameraRayProjection::reset, src/slim/scene/camera.h:69-83)
            d = rot[0] * (x + 0.5 - 32.0) + rot[1] * (18.0 - y - 0.5) + rot[2] * (18.0 * focal)
            d /= np.linalg.norm(d)
            P = cam + d * float(t[y, x])
            N = np.array([0.0, 1.0, 0.0])
            L = lp - P
            Ld2 = L @ L
            L /= np.sqrt(Ld2)
            NdotL = max(L @ N, 0.0)
            R = d - 2.0 * (d @ N) * N
            spec = max(R @ L, 0.0)
            c = NdotL * inten / Ld2
            rad = albedo * rough / np.pi * c + (f0 * spec ** 4 * (1 - rough) * c if spec > 0 else 0)
            x0 = np.maximum(rad - 0.004, 0.0)
            tm = (6.2 * x0 * x0 + 0.5 * x0) / (6.2 * x0 * x0 + 1.7 * x0 + 0.06)
            assert np.allclose(col[y, x], tm, atol=5e-5), (x, y, col[y, x], tm)
            checked += 1
    assert checked > 20
