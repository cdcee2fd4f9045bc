"""Shared helpers for the tests (spec round trips, mesh-spec construction from asset files)."""
import glob
import hashlib
import os
import tempfile

import numpy as np

from slimtracincpp_amd import scenes
from slimtracincpp_amd.specfile import spec_bytes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_cases():
    return sorted(os.path.basename(p)[4:-4] for p in glob.glob(os.path.join(GOLDEN, "ref_*.npz")))


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, f"ref_{name}.npz"), allow_pickle=False))


def golden_spec_path(name, g, tmpdir):
    """Writes the golden case's spec to tmpdir; regenerates it from the seeded generator when only its digest
    is stored, and checks the digest."""
    path = os.path.join(tmpdir, f"{name}.spec")
    if "spec" in g:
        data = g["spec"].tobytes()
    else:
        from tests.golden.make_golden import CASES
        data = spec_bytes(CASES[name]())
        assert hashlib.sha256(data).hexdigest() == str(g["spec__sha256"]), "seeded generator drifted"
    with open(path, "wb") as f:
        f.write(data)
    return path


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def mesh_spec_from_asset(m: dict, leaf: int) -> scenes.SceneSpec:
    """A one-geometry scene around a mesh read from a reference .mesh file (for builder comparisons)."""
    ms = scenes.MeshSpec(positions=m["positions"].copy(), position_indices=m["position_indices"].copy(),
                         normals=m.get("normals"), normal_indices=m.get("normal_indices"),
                         uvs=m.get("uvs"), uv_indices=m.get("uv_indices"), leaf_size=leaf)
    f, u = scenes.material()
    g = scenes._geo(scenes.GEO_MESH, (0, 0, 0), (0, 0, 0), (1, 1, 1), 0, 0)
    return scenes.SceneSpec(name="asset", width=8, height=8, camera_euler=np.zeros(3, np.float32),
                            camera_position=np.array([0, 0, -10], np.float32), focal_length=np.float32(2.0),
                            geo_f=g[0][None], geo_u=g[1][None], mat_f=f[None], mat_u=u[None],
                            light_f=np.zeros((0, 7), np.float32), light_u=np.zeros(0, np.uint32), meshes=[ms])


def shipped_sky_scene(width: int = 96, height: int = 64, seed: int = 7) -> scenes.SceneSpec:
    """A small mesh + PBR scene lit by the reference's shipped cube map (bolonga_irradiance.texture) as sky colour,
    radiance and irradiance, so texture sampling of cube maps is exercised on data the reference's converter wrote."""
    from slimtracincpp_amd import assets
    t = assets.read_texture(os.path.join(GOLDEN, "bolonga_irradiance.texture"))
    cube = scenes.TextureSpec(width=t["width"], height=t["height"], flags=t["flags"],
                              mips=[q for _, _, q in t["mips"]], mip_dims=[(w, h) for w, h, _ in t["mips"]])
    spec = scenes.mesh_pbr_scene(seed, width, height, level=3, texture_size=32, sky_size=8, irradiance_size=4)
    spec.textures = list(spec.textures) + [cube]
    k = len(spec.textures) - 1
    spec.skybox_color = spec.skybox_radiance = spec.skybox_irradiance = k
    return spec
