"""Seeded synthetic scenes for the per-pixel ray-tracing path.

The reference ships no random generator and no benchmark scenes, so the
scenes used for parity tests and for measurement are defined here, in the
reference's own vocabulary (geometries with Euler orientations, materials,
point lights, meshes given as vertex/index arrays, textures given as mip
chains of RGB-u8 texel quads).  A `SceneSpec` is the *input* of the path: the
product's host code, the CPU oracle and the compiled reference each derive
everything else (quaternions, AABBs, BVHs, triangle records, the camera ray
projection) from the same spec on their own.

Random draws come from PCG32 (O'Neill, "PCG: A Family of Simple Fast
Space-Efficient Statistically Good Algorithms", 2014) seeded through
SplitMix64 (Steele et al., 2014).  Every uniform draw is
``(next_u32() >> 8) * 2**-24`` mapped affinely and rounded to float32; the draw
order is the order of the statements in each builder below.

Enum values follow the reference: GeometryType (src/slim/core/base.h:399-410),
BRDFType (base.h:624-629), RenderMode (base.h:517-524), material flags
(base.h:142-146), geometry flags (base.h:148-150), image flags (base.h:1283-1295).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# --- enums (values as in the reference) -------------------------------------
GEO_NONE, GEO_MESH, GEO_GRID, GEO_BOX, GEO_CURVE, GEO_TET, GEO_QUAD, GEO_SPHERE = range(8)
BRDF_LAMBERT, BRDF_BLINN, BRDF_PHONG, BRDF_COOK_TORRANCE = range(4)
MODE_NORMALS, MODE_NORMAL_MAP, MODE_BEAUTY, MODE_DEPTH, MODE_MIP_LEVEL, MODE_UVS = range(6)

MAT_REFLECTIVE = 1
MAT_REFRACTIVE = 2
MAT_EMISSIVE = 4
MAT_ALBEDO_MAP = 16
MAT_NORMAL_MAP = 32

GEO_VISIBLE = 1
GEO_SHADOWING = 2
GEO_TRANSPARENT = 4

LIGHT_DIRECTIONAL = 1
LIGHT_SHADOWING = 2

IMG_LINEAR = 1 << 1
IMG_MIPMAP = 1 << 4
IMG_WRAP = 1 << 6
IMG_NORMAL = 1 << 7
IMG_CUBEMAP = 1 << 8
# texel storage (texture flags bits 24-25; include/slimrt.h SLIMRT_TEXELS_*): the reference's texel quads, or a
# bordered grid of RGBA8 or RGBA16F texels
TEXELS_QUADS = 0
TEXELS_RGBA8 = 1 << 24
TEXELS_RGBA16F = 2 << 24
TEXELS_MASK = 3 << 24

IOR_AIR = np.float32(1.0003)

MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


class PCG32:
    """pcg32 (XSH-RR, 64-bit state), seeded as pcg32_srandom(splitmix64(seed), splitmix64(seed + 1))."""

    MULT = 6364136223846793005

    def __init__(self, seed: int):
        initstate = splitmix64(seed)
        initseq = splitmix64((seed + 1) & MASK64)
        self.state = 0
        self.inc = ((initseq << 1) | 1) & MASK64
        self.next_u32()
        self.state = (self.state + initstate) & MASK64
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * self.MULT + self.inc) & MASK64
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> np.float32:
        u = (self.next_u32() >> 8) * (1.0 / 16777216.0)
        return np.float32(lo + (hi - lo) * u)

    def uniforms(self, n: int, lo: float, hi: float) -> np.ndarray:
        return np.array([self.uniform(lo, hi) for _ in range(n)], dtype=np.float32)


# --- spec containers -----------------------------------------------------------
@dataclass
class MeshSpec:
    positions: np.ndarray          # (V,3) f32
    position_indices: np.ndarray   # (T,3) u32
    normals: Optional[np.ndarray] = None        # (N,3) f32
    normal_indices: Optional[np.ndarray] = None  # (T,3) u32
    uvs: Optional[np.ndarray] = None            # (U,2) f32
    uv_indices: Optional[np.ndarray] = None     # (T,3) u32
    leaf_size: int = 2             # max triangles per BVH leaf
    edges: Optional[np.ndarray] = None          # (E,2) u32 vertex pairs (EdgeVertexIndices), for wireframe overlays

    @property
    def triangle_count(self) -> int:
        return int(self.position_indices.shape[0])


def with_mesh_edges(spec: "SceneSpec") -> "SceneSpec":
    """The spec with every mesh's edge list filled in (triangle_edges), for wireframe overlays."""
    for m in spec.meshes:
        m.edges = triangle_edges(m.position_indices)
    return spec


def triangle_edges(position_indices: np.ndarray) -> np.ndarray:
    """The distinct edges of a triangle list as (E, 2) u32 vertex pairs, in order of first appearance (each
    triangle's edges v1-v2, v2-v3, v3-v1), every pair ordered as it first appears."""
    tri = np.asarray(position_indices, np.uint32).reshape(-1, 3)
    pairs = np.stack([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], axis=1).reshape(-1, 2)
    key = np.sort(pairs, axis=1).astype(np.uint64)
    key = key[:, 0] << np.uint64(32) | key[:, 1]
    _, first = np.unique(key, return_index=True)
    return np.ascontiguousarray(pairs[np.sort(first)], np.uint32)


@dataclass
class TextureSpec:
    width: int
    height: int
    flags: int
    mips: List[np.ndarray]         # each (h+1, w+1, 12) u8 texel quads: R{TL,TR,BL,BR} G{..} B{..}, or with
                                   # TEXELS_RGBA8 / TEXELS_RGBA16F in flags (h+2, w+2, 4) u8 / float16 texels
    mip_dims: List[tuple]          # (w, h) per mip

    @property
    def texel_storage(self) -> int:
        return self.flags & TEXELS_MASK


@dataclass
class SceneSpec:
    name: str
    width: int
    height: int
    camera_euler: np.ndarray       # (3,) f32 radians, Orientation<mat3> order (z, then x, then y)
    camera_position: np.ndarray    # (3,) f32
    focal_length: np.float32
    max_depth: int = 3
    render_mode: int = MODE_BEAUTY
    skybox_color: int = -1
    skybox_radiance: int = -1
    skybox_irradiance: int = -1
    ssaa: bool = False
    geo_f: np.ndarray = None       # (G,9) f32: euler xyz | position xyz | scale xyz
    geo_u: np.ndarray = None       # (G,4) u32: type | material_id | id (mesh index) | flags
    mat_f: np.ndarray = None       # (M,16) f32: see MAT_F_* below
    mat_u: np.ndarray = None       # (M,5) u32: flags | texture_count | texture_id0 | texture_id1 | brdf
    light_f: np.ndarray = None     # (L,7) f32: color rgb | position xyz | intensity
    light_u: np.ndarray = None     # (L,) u32 flags
    meshes: List[MeshSpec] = field(default_factory=list)
    textures: List[TextureSpec] = field(default_factory=list)

    @property
    def geometry_count(self) -> int:
        return int(self.geo_u.shape[0])


# column layout of SceneSpec.mat_f
MAT_F_ALBEDO = slice(0, 3)
MAT_F_ROUGHNESS = 3
MAT_F_NORMAL_MAGNITUDE = 4
MAT_F_UV_REPEAT = slice(5, 7)
MAT_F_EMISSION = slice(7, 10)
MAT_F_IOR = 10
MAT_F_REFLECTIVITY = slice(11, 14)
MAT_F_METALNESS = 14


def material(albedo=(1.0, 1.0, 1.0), roughness=1.0, flags=0, texture_ids=(), brdf=BRDF_COOK_TORRANCE,
             normal_magnitude=1.0, uv_repeat=(1.0, 1.0), emission=(0.0, 0.0, 0.0), ior=1.0,
             reflectivity=(0.04, 0.04, 0.04), metalness=0.0):
    """One material row pair (f, u); defaults are those of `Material` (src/slim/scene/material.h:93-107)."""
    f = np.zeros(16, np.float32)
    f[MAT_F_ALBEDO] = albedo
    f[MAT_F_ROUGHNESS] = roughness
    f[MAT_F_NORMAL_MAGNITUDE] = normal_magnitude
    f[MAT_F_UV_REPEAT] = uv_repeat
    f[MAT_F_EMISSION] = emission
    f[MAT_F_IOR] = ior
    f[MAT_F_REFLECTIVITY] = reflectivity
    f[MAT_F_METALNESS] = metalness
    ids = list(texture_ids) + [0, 0]
    u = np.array([flags, len(texture_ids), ids[0], ids[1], brdf], np.uint32)
    return f, u


def _geo(gtype, euler, position, scale, material_id, mesh_id=0, flags=GEO_VISIBLE | GEO_SHADOWING):
    f = np.concatenate([np.asarray(euler, np.float32), np.asarray(position, np.float32),
                        np.asarray(scale, np.float32)]).astype(np.float32)
    u = np.array([gtype, material_id, mesh_id, flags], np.uint32)
    return f, u


def _stack(rows):
    return np.stack([r for r in rows]).astype(rows[0].dtype)


FOV60_FOCAL = np.float32(1.0 / math.tan(math.radians(30.0)))


def look_at_euler(position, target=(0.0, 0.0, 0.0)):
    """Euler (pitch x, yaw y, 0) for a camera looking at target.

    The camera's forward axis is the third row of Rx(pitch) * Ry(yaw) (src/slim/math/mat3.h:57-77 composed as in
    src/slim/core/base.h:612-620): (-cos p sin y, sin p, cos p cos y).  Pitch is negative when looking down, as in
    the reference demos (src/examples/01_Lights.cpp:25).
    """
    d = np.asarray(target, np.float64) - np.asarray(position, np.float64)
    yaw = math.atan2(-d[0], d[2])
    pitch = math.atan2(d[1], math.hypot(d[0], d[2]))
    return np.array([pitch, yaw, 0.0], np.float32)


# --- scene 1: implicit primitives --------------------------------------------------
def implicit_scene(seed: int = 0, width: int = 320, height: int = 180, max_depth: int = 2,
                   transparent: bool = False, render_mode: int = MODE_BEAUTY) -> SceneSpec:
    """The plain CPU-runnable scene: 20x20 floor quad, 4 spheres, 2 boxes, 1 tetrahedron, 2 point lights.

    Draw order: spheres (radius, position xyz, euler xyz) x4; boxes (half-extents xyz, position, euler) x2;
    tetrahedron (scale, position, euler); lights (intensity, position xyz) x2; then one material per
    geometry in geometry order (floor first): brdf coin, albedo rgb, shininess, roughness.
    Geometry order: floor, sphere0..3, box0..1, tet.  Sphere 0 is the mirror.
    `transparent=True` additionally marks boxes, the tetrahedron and spheres 2-3 as UV-checkerboard
    transparent (GEOMETRY_IS_TRANSPARENT); the configuration itself has no transparent geometry.
    """
    rng = PCG32(seed)
    geos = []
    two_pi = 2.0 * math.pi
    geos.append(_geo(GEO_QUAD, (0, 0, 0), (0, 0, 0), (10, 1, 10), 0))

    def placement():
        pos = (rng.uniform(-4, 4), rng.uniform(0.5, 3), rng.uniform(-4, 4))
        eul = (rng.uniform(0, two_pi), rng.uniform(0, two_pi), rng.uniform(0, two_pi))
        return pos, eul

    tflag = GEO_TRANSPARENT if transparent else 0
    for i in range(4):
        r = rng.uniform(0.5, 1.5)
        pos, eul = placement()
        fl = GEO_VISIBLE | GEO_SHADOWING | (tflag if i >= 2 else 0)
        geos.append(_geo(GEO_SPHERE, eul, pos, (r, r, r), len(geos), flags=fl))
    for i in range(2):
        h = (rng.uniform(0.4, 1.2), rng.uniform(0.4, 1.2), rng.uniform(0.4, 1.2))
        pos, eul = placement()
        geos.append(_geo(GEO_BOX, eul, pos, h, len(geos), flags=GEO_VISIBLE | GEO_SHADOWING | tflag))
    s = rng.uniform(0.8, 1.6)
    pos, eul = placement()
    geos.append(_geo(GEO_TET, eul, pos, (s, s, s), len(geos), flags=GEO_VISIBLE | GEO_SHADOWING | tflag))

    lights_f, lights_u = [], []
    for _ in range(2):
        intensity = rng.uniform(5, 20)
        p = (rng.uniform(-5, 5), rng.uniform(4, 8), rng.uniform(-5, 5))
        lights_f.append(np.array([1, 1, 1, p[0], p[1], p[2], intensity], np.float32))
        lights_u.append(LIGHT_SHADOWING)

    mats_f, mats_u = [], []
    for gi in range(len(geos)):
        coin = rng.uniform()
        albedo = (rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9))
        _shininess = rng.uniform(8, 128)   # drawn for the configuration; the reference has no such field
        rough = rng.uniform(0.3, 0.9)
        if gi == 0:
            f, u = material(albedo, 1.0, brdf=BRDF_LAMBERT)
        elif gi == 1:   # mirror sphere
            f, u = material(albedo, rough, flags=MAT_REFLECTIVE, brdf=BRDF_BLINN, reflectivity=(0.9, 0.9, 0.9))
        elif coin < 0.5:
            f, u = material(albedo, 1.0, brdf=BRDF_LAMBERT)
        else:
            f, u = material(albedo, rough, brdf=BRDF_BLINN)
        mats_f.append(f)
        mats_u.append(u)

    cam_pos = np.array([0.0, 2.0, -8.0], np.float32)
    cam_euler = np.array([-math.atan2(2.0, 8.0), 0.0, 0.0], np.float32)
    return SceneSpec(
        name=f"implicit_s{seed}_{width}x{height}", width=width, height=height,
        camera_euler=cam_euler, camera_position=cam_pos, focal_length=FOV60_FOCAL,
        max_depth=max_depth, render_mode=render_mode,
        geo_f=_stack([g[0] for g in geos]), geo_u=_stack([g[1] for g in geos]),
        mat_f=_stack(mats_f), mat_u=_stack(mats_u),
        light_f=_stack(lights_f), light_u=np.array(lights_u, np.uint32))


# --- procedural meshes ---------------------------------------------------------------
def _hash3(ix, iy, iz, seed):
    """Integer lattice hash -> float64 in [-1, 1) (SplitMix64 finaliser on a packed key)."""
    with np.errstate(over="ignore"):
        k = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ \
            (iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)) ^ \
            (iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)) ^ np.uint64(splitmix64(seed))
        k = (k ^ (k >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        k = (k ^ (k >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        k = k ^ (k >> np.uint64(31))
    return (k >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def value_noise3(p: np.ndarray, seed: int, octaves: int = 4, base_frequency: float = 2.0) -> np.ndarray:
    """Seeded 3D value noise: trilinear with smoothstep weights, `octaves` octaves of doubling frequency and
    halving amplitude, normalised to [-1, 1].  p: (N, 3) float64.  The lattice hash is evaluated once per lattice
    point of the octave's bounding box and gathered (the same values as hashing every corner of every point)."""
    total = np.zeros(p.shape[0])
    amp, norm, freq = 1.0, 0.0, base_frequency
    for o in range(octaves):
        q = p * freq
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        w = f * f * (3.0 - 2.0 * f)
        lo = i0.min(axis=0)
        rel = i0 - lo
        dims = rel.max(axis=0) + 2
        gx, gy, gz = np.meshgrid(*[np.arange(d, dtype=np.int64) + l for d, l in zip(dims, lo)], indexing="ij")
        table = _hash3(gx.ravel(), gy.ravel(), gz.ravel(), seed + 7919 * o)
        base = (rel[:, 0] * dims[1] + rel[:, 1]) * dims[2] + rel[:, 2]
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            wx = w[:, 0] if dx else 1 - w[:, 0]
            for dy in (0, 1):
                wy = w[:, 1] if dy else 1 - w[:, 1]
                for dz in (0, 1):
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc += wx * wy * wz * table[base + (dx * dims[1] + dy) * dims[2] + dz]
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def icosphere(level: int):
    """Unit icosphere: 20 * 4**level triangles, 10 * 4**level + 2 vertices (float64 positions, int64 faces)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        n = f.shape[0]
        a, b, c = inv[:n] + base, inv[n:2 * n] + base, inv[2 * n:] + base
        f = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
                            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return v, f


def displaced_sphere_mesh(level: int = 6, amplitude: float = 0.15, seed: int = 0, leaf_size: int = 4,
                          uv_sphere: tuple | None = None) -> MeshSpec:
    """Icosphere (or, with uv_sphere=(rings, segments), a latitude/longitude sphere) displaced along its normal by
    `amplitude` * 4-octave value noise, with smooth (area-weighted) vertex normals and spherical UVs.
    Positions are recentred on their AABB centre as the reference's mesh converter does (src/obj2mesh.cpp:215-230)."""
    if uv_sphere is None:
        d, f = icosphere(level)
    else:
        rings, segs = uv_sphere
        th = (np.arange(1, rings) / rings) * math.pi
        ph = np.arange(segs) / segs * 2 * math.pi
        TH, PH = np.meshgrid(th, ph, indexing="ij")
        d = np.stack([np.sin(TH) * np.cos(PH), np.cos(TH), np.sin(TH) * np.sin(PH)], -1).reshape(-1, 3)
        d = np.concatenate([[[0, 1, 0]], d, [[0, -1, 0]]])
        idx = lambda r, s: 1 + r * segs + (s % segs)
        faces = []
        for s in range(segs):
            faces.append([0, idx(0, s + 1), idx(0, s)])
            faces.append([d.shape[0] - 1, idx(rings - 2, s), idx(rings - 2, s + 1)])
        r = np.arange(rings - 2)[:, None]
        s = np.arange(segs)[None, :]
        a, b, c, e = idx(r, s), idx(r, s + 1), idx(r + 1, s), idx(r + 1, s + 1)
        quads = np.stack([np.stack([a, b, c], -1), np.stack([b, e, c], -1)], 2).reshape(-1, 3)
        f = np.concatenate([np.array(faces), quads]).astype(np.int64)
    disp = 1.0 + amplitude * value_noise3(d, seed)
    p = d * disp[:, None]
    fn = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])   # counter-clockwise faces: outward
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    uv = np.stack([0.5 + np.arctan2(d[:, 2], d[:, 0]) / (2 * math.pi), 0.5 - np.arcsin(np.clip(d[:, 1], -1, 1)) / math.pi], 1)
    pos = p.astype(np.float32)
    mn, mx = pos.min(0), pos.max(0)
    centroid = (mn + mx) * np.float32(0.5)
    if np.any(centroid != 0):
        pos = (pos - centroid).astype(np.float32)
    # The reference's triangle frame is X = v3 - v1, Y = v2 - v1, normal = X x Y (src/slim/scene/bvh_builder.h:315-319),
    # so faces are stored clockwise as seen from outside for that normal to point outwards.
    fi = f[:, [0, 2, 1]].astype(np.uint32)
    return MeshSpec(positions=pos, position_indices=fi, normals=n.astype(np.float32), normal_indices=fi.copy(),
                    uvs=uv.astype(np.float32), uv_indices=fi.copy(), leaf_size=leaf_size)


# --- textures in the reference's texel-quad layout ------------------------------------
def bake_texel_quads(texels: np.ndarray, wrap: bool) -> np.ndarray:
    """float texels (h, w, 3) in [0, 1] -> (h+1, w+1, 12) u8 TexelQuads (R{TL,TR,BL,BR} G{..} B{..}).
    Quad (x, y) holds texels (x-1, y-1), (x, y-1), (x-1, y), (x, y); out-of-range indices wrap (as a torus) or
    clamp.  The converter (TextureMipLoader::load, src/bmp2texture.cpp:36-95; slimrt_host_bake_texture) gives the
    same quads except, with wrap, the right border column's TR and BR corners, which its later right-edge writes
    leave at the last column's texels.  Both renderers read the same quads, so this is a choice of input, not a
    parity difference.  Quantisation truncates x * 255 (src/bmp2texture.cpp:327-357)."""
    return quantize_quads(pad_texels(texels, wrap))


def quantize_quads(pad: np.ndarray) -> np.ndarray:
    """Padded texels (h+2, w+2, 3) -> (h+1, w+1, 12) u8 quads.  Each texel is quantised once (x * 255 in float32,
    truncated) and the quads gather the quantised corners."""
    u8 = (pad.astype(np.float32) * np.float32(255.0)).astype(np.float32).astype(np.uint8)
    h1, w1 = u8.shape[0] - 1, u8.shape[1] - 1
    q = np.empty((h1, w1, 3, 4), np.uint8)             # R{TL,TR,BL,BR} G{..} B{..}
    q[..., 0] = u8[:-1, :-1]
    q[..., 1] = u8[:-1, 1:]
    q[..., 2] = u8[1:, :-1]
    q[..., 3] = u8[1:, 1:]
    return q.reshape(h1, w1, 12)


def quantize_texels(pad: np.ndarray) -> np.ndarray:
    """Padded texels (h+2, w+2, 3) -> (h+2, w+2, 4) RGBA8 (A = 255), quantised as quantize_quads does, so the
    grid holds exactly the corners of the quads quantize_quads makes from the same texels."""
    u8 = (pad.astype(np.float32) * np.float32(255.0)).astype(np.float32).astype(np.uint8)
    return np.concatenate([u8, np.full(u8.shape[:2] + (1,), 255, np.uint8)], -1)


def half_texels(pad: np.ndarray) -> np.ndarray:
    """Padded float texels (h+2, w+2, 3) -> (h+2, w+2, 4) float16 (A = 1), each value rounded to the nearest
    half (through float32)."""
    f = pad.astype(np.float32)
    return np.concatenate([f, np.ones(f.shape[:2] + (1,), np.float32)], -1).astype(np.float16)


def _store(pad: np.ndarray, texel_storage: int) -> np.ndarray:
    if texel_storage == TEXELS_RGBA8:
        return quantize_texels(np.clip(pad, 0, 1))
    if texel_storage == TEXELS_RGBA16F:
        return half_texels(pad)
    return quantize_quads(np.clip(pad, 0, 1))


def texels_from_quads(q: np.ndarray, texel_storage: int):
    """(h+1, w+1, 12) quads -> (bordered texel grid, number of quads whose corners differ from it); the numpy
    form of slimrt_host_texels_from_quads.  Grid texel (x, y) is quad (x, y)'s top-left corner; the last column
    and row are the last quads' right and bottom corners."""
    c = q.reshape(q.shape[0], q.shape[1], 3, 4)        # [y, x, channel, corner TL TR BL BR]
    h1, w1 = c.shape[:2]
    g = np.empty((h1 + 1, w1 + 1, 3), np.uint8)
    g[:h1, :w1] = c[..., 0]
    g[:h1, w1] = c[:, -1, :, 1]
    g[h1, :w1] = c[-1, :, :, 2]
    g[h1, w1] = c[-1, -1, :, 3]
    back = np.stack([g[:-1, :-1], g[:-1, 1:], g[1:, :-1], g[1:, 1:]], -1)
    mismatches = int(np.count_nonzero(np.any(back != c, axis=(2, 3))))
    if texel_storage == TEXELS_RGBA8:
        return np.concatenate([g, np.full(g.shape[:2] + (1,), 255, np.uint8)], -1), mismatches
    f = g.astype(np.float32) / np.float32(255.0)
    return np.concatenate([f, np.ones(f.shape[:2] + (1,), np.float32)], -1).astype(np.float16), mismatches


def convert_texture(t: TextureSpec, texel_storage: int) -> TextureSpec:
    """A quad texture in texel storage (RGBA8: the same samples when no quad mismatches its grid)."""
    assert t.texel_storage == TEXELS_QUADS
    mips = [texels_from_quads(q, texel_storage)[0] for q in t.mips]
    return TextureSpec(width=t.width, height=t.height, flags=t.flags | texel_storage, mips=mips,
                       mip_dims=list(t.mip_dims))


def mip_chain(texels: np.ndarray):
    """Box-filtered chain down to 4 px as TextureMipLoader/loadMips do (src/bmp2texture.cpp:201-237, 289-299)."""
    mips = [texels.astype(np.float32)]
    while mips[-1].shape[0] > 4 and mips[-1].shape[1] > 4:
        t = mips[-1]
        a, b, c, d = t[0::2, 0::2], t[0::2, 1::2], t[1::2, 0::2], t[1::2, 1::2]
        mips.append((np.float32(0.25) * (((a + b) + c) + d)).astype(np.float32))
    return mips


def pad_texels(texels: np.ndarray, wrap: bool) -> np.ndarray:
    """(h, w, 3) -> (h+2, w+2, 3): a one-texel border from the opposite edge (wrap, as a torus) or the edge."""
    h, w, _ = texels.shape
    xs = np.arange(-1, w + 1)
    ys = np.arange(-1, h + 1)
    xs = xs % w if wrap else np.clip(xs, 0, w - 1)
    ys = ys % h if wrap else np.clip(ys, 0, h - 1)
    return texels[ys][:, xs].astype(np.float32)


def texture_from_texels(texels: np.ndarray, mipmap: bool = True, wrap: bool = True,
                        texel_storage: int = TEXELS_QUADS) -> TextureSpec:
    chain = mip_chain(texels) if mipmap else [texels.astype(np.float32)]
    h, w, _ = texels.shape
    flags = (IMG_MIPMAP if mipmap else 0) | (IMG_WRAP if wrap else 0) | texel_storage
    if texel_storage == TEXELS_QUADS:
        mips = [bake_texel_quads(m, wrap) for m in chain]
    else:
        mips = [_store(pad_texels(m, wrap), texel_storage) for m in chain]
    return TextureSpec(width=w, height=h, flags=flags, mips=mips, mip_dims=[(m.shape[1], m.shape[0]) for m in chain])


def checker_noise_texels(size: int = 1024, squares: int = 8, seed: int = 0) -> np.ndarray:
    """RGBA8 checkerboard of two seeded colours modulated by 2D value noise; returns RGB floats (u8 / 255)."""
    rng = PCG32(seed)
    c0 = np.array([rng.uniform(0.55, 0.95) for _ in range(3)])
    c1 = np.array([rng.uniform(0.05, 0.45) for _ in range(3)])
    y, x = np.mgrid[0:size, 0:size]
    check = ((x * squares // size) + (y * squares // size)) % 2
    p = np.stack([x.ravel() / size * 16.0, y.ravel() / size * 16.0, np.zeros(size * size)], 1)
    n = value_noise3(p, seed + 11, octaves=4, base_frequency=1.0).reshape(size, size)
    base = np.where(check[..., None] == 1, c0, c1)
    rgb = np.clip(base * (0.8 + 0.2 * n[..., None]), 0, 1)
    rgba8 = np.concatenate([(rgb * 255).astype(np.uint8), np.full((size, size, 1), 255, np.uint8)], -1)
    return rgba8[..., :3].astype(np.float32) / np.float32(255.0)


# Cube faces in sampleCube's layout (src/slim/core/texture.h:72-116): the 4-face strip L(-X) F(+Z) R(+X) B(-Z),
# then top (+Y) and bottom (-Y).  Direction of face texel (s, t) in [0,1]^2, inverted from that mapping:
_CUBE_DIRS = {
    "L": lambda s, t: np.stack([-np.ones_like(s), 1 - 2 * t, 2 * s - 1], -1),
    "F": lambda s, t: np.stack([2 * s - 1, 1 - 2 * t, np.ones_like(s)], -1),
    "R": lambda s, t: np.stack([np.ones_like(s), 1 - 2 * t, 1 - 2 * s], -1),
    "B": lambda s, t: np.stack([1 - 2 * s, 1 - 2 * t, -np.ones_like(s)], -1),
    "T": lambda s, t: np.stack([2 * s - 1, np.ones_like(s), 2 * t - 1], -1),
    "D": lambda s, t: np.stack([2 * s - 1, -np.ones_like(s), 1 - 2 * t], -1),
}


def _face_dirs(face: str, n: int, pad: int = 1):
    idx = (np.arange(-pad, n + pad) + 0.5) / n
    t, s = np.meshgrid(idx, idx, indexing="ij")
    d = _CUBE_DIRS[face](s, t)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def cube_texture(fn, n: int, texel_storage: int = TEXELS_QUADS) -> TextureSpec:
    """Cube map of face size n from a direction -> RGB function, as 3 mips: strip (4n x n), top, bottom.
    Border quads take the function's value one texel beyond each face edge, which is the seamless result that
    the reference converter's border stitching approximates (src/bmp2texture.cpp:97-198).  In RGBA16F storage
    the values are kept as they are (above 1 for an HDR function); quads and RGBA8 clip them to [0, 1]."""
    # strip: rows -1..n, columns -1..4n (one padding column each side, taken from the L and B extensions)
    left_pad = fn(_face_dirs("L", n))[:, :1]
    body = np.concatenate([fn(_face_dirs(f, n))[:, 1:-1] for f in "LFRB"], axis=1)
    right_pad = fn(_face_dirs("B", n))[:, -1:]
    strip = np.concatenate([left_pad, body, right_pad], axis=1).astype(np.float32)
    top = fn(_face_dirs("T", n)).astype(np.float32)
    bottom = fn(_face_dirs("D", n)).astype(np.float32)
    mips = [_store(strip, texel_storage), _store(top, texel_storage), _store(bottom, texel_storage)]
    return TextureSpec(width=6 * n, height=n, flags=IMG_CUBEMAP | IMG_LINEAR | texel_storage, mips=mips,
                       mip_dims=[(4 * n, n), (n, n), (n, n)])


def sky_function(seed: int = 0, spots: int = 3, hdr: bool = False):
    """Seeded sky: vertical gradient (ground, horizon, zenith) plus `spots` Gaussian hot spots; clipped to [0, 1]
    unless hdr (for half-float storage: the hot spots then keep their peaks above 1)."""
    rng = PCG32(seed)
    centers, colors, sharp = [], [], []
    for _ in range(spots):
        th = rng.uniform(0.15, 1.3)          # elevation above the horizon, radians
        ph = rng.uniform(0, 2 * math.pi)
        centers.append([math.cos(th) * math.cos(ph), math.sin(th), math.cos(th) * math.sin(ph)])
        colors.append([rng.uniform(0.5, 1.0), rng.uniform(0.4, 0.9), rng.uniform(0.3, 0.8)])
        sharp.append(rng.uniform(20, 120))
    centers, colors, sharp = np.array(centers), np.array(colors), np.array(sharp)
    ground, horizon, zenith = np.array([0.25, 0.22, 0.2]), np.array([0.75, 0.8, 0.85]), np.array([0.2, 0.35, 0.7])

    def fn(d):
        y = d[..., 1:2]
        up = np.clip(y, 0, 1)
        col = np.where(y >= 0, horizon * (1 - up) + zenith * up, horizon * np.clip(1 + 4 * y, 0, 1) +
                       ground * (1 - np.clip(1 + 4 * y, 0, 1)))
        for c, k, s in zip(centers, colors, sharp):
            col = col + k * np.exp(s * (np.sum(d * c, -1, keepdims=True) - 1.0))
        return col * 0.8 if hdr else np.clip(col * 0.8, 0, 1)
    return fn


def irradiance_function(radiance_fn, samples: int = 16):
    """Cosine-weighted irradiance (normalised so a constant sky maps to itself) by quadrature over cube texels."""
    dirs, weights = [], []
    for f in "LFRBTD":
        d = _face_dirs(f, samples, pad=0).reshape(-1, 3)
        idx = (np.arange(samples) + 0.5) / samples * 2 - 1
        a, b = np.meshgrid(idx, idx, indexing="ij")
        w = (2.0 / samples) ** 2 / (1 + a * a + b * b) ** 1.5
        dirs.append(d)
        weights.append(w.reshape(-1))
    dirs, weights = np.concatenate(dirs), np.concatenate(weights)
    L = radiance_fn(dirs)

    def fn(n):
        shp = n.shape[:-1]
        nn = n.reshape(-1, 3)
        cos = np.clip(nn @ dirs.T, 0, None) * weights[None, :]
        return ((cos @ L) / math.pi).reshape(*shp, 3)
    return fn


# --- scene 2: meshes + PBR + IBL ---------------------------------------------------------
def mesh_pbr_scene(seed: int = 0, width: int = 1920, height: int = 1080, max_depth: int = 4, level: int = 6,
                   texture_size: int = 1024, sky_size: int = 256, irradiance_size: int = 32,
                   render_mode: int = MODE_BEAUTY, leaf_size: int = 4, texel_storage: bool = False) -> SceneSpec:
    """Three instances of a displaced icosphere (81,920 triangles at level 6) on a textured floor quad, lit by
    cube-map IBL only.  Instance 0 is glass (IOR 1.5), instance 1 a mirror, instance 2 textured GGX.
    Draw order: per instance (scale, position x z, euler xyz, roughness, metallic coin), then floor roughness.
    Textures: 0 = checker+noise albedo (mip-mapped, wrapped), 1 = sky radiance cube, 2 = irradiance cube; in the
    reference's texel quads, or with texel_storage the albedo as RGBA8 texels and the cube faces as RGBA16F
    texels of an HDR sky."""
    rng = PCG32(seed)
    mesh = displaced_sphere_mesh(level, 0.15, seed, leaf_size)
    geos, mats_f, mats_u = [], [], []
    geos.append(_geo(GEO_QUAD, (0, 0, 0), (0, 0, 0), (10, 1, 10), 0))
    slots = [(-2.6, 0.0), (0.0, 1.2), (2.6, -0.4)]
    inst = []
    for i in range(3):
        s = rng.uniform(0.7, 1.5)
        px, pz = rng.uniform(-0.4, 0.4) + slots[i][0], rng.uniform(-0.4, 0.4) + slots[i][1]
        eul = (rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi))
        rough = rng.uniform(0.05, 0.9)
        metal = 1.0 if rng.uniform() < 0.5 else 0.0
        inst.append((s, px, pz, eul, rough, metal))
    floor_rough = rng.uniform(0.05, 0.9)
    # floor material
    f, u = material((1, 1, 1), floor_rough, flags=MAT_ALBEDO_MAP, texture_ids=(0,), brdf=BRDF_COOK_TORRANCE,
                    uv_repeat=(4, 4), reflectivity=(0.04, 0.04, 0.04), metalness=0.0)
    mats_f.append(f); mats_u.append(u)
    for i, (s, px, pz, eul, rough, metal) in enumerate(inst):
        fl = GEO_VISIBLE | GEO_SHADOWING
        if i == 0:      # glass
            f, u = material((0.05, 0.05, 0.05), rough * 0.2, flags=MAT_REFRACTIVE, ior=1.5,
                            reflectivity=(0.04, 0.04, 0.04))
            fl = GEO_VISIBLE
        elif i == 1:    # mirror
            f, u = material((0.0, 0.0, 0.0), 0.02, flags=MAT_REFLECTIVE, reflectivity=(0.91, 0.92, 0.92), metalness=1.0)
        else:
            base = np.array([0.9, 0.7, 0.4], np.float32)
            f0 = 0.04 * (1 - metal) + base * metal
            f, u = material(base, rough, flags=MAT_ALBEDO_MAP, texture_ids=(0,), reflectivity=tuple(f0),
                            metalness=metal)
        mats_f.append(f); mats_u.append(u)
        geos.append(_geo(GEO_MESH, eul, (px, s * 1.15, pz), (s, s, s), len(geos), 0, fl))
    sky = sky_function(seed + 101, hdr=texel_storage)
    albedo_storage, cube_storage = (TEXELS_RGBA8, TEXELS_RGBA16F) if texel_storage else (TEXELS_QUADS, TEXELS_QUADS)
    tex = [texture_from_texels(checker_noise_texels(texture_size, 8, seed + 202), True, True, albedo_storage),
           cube_texture(sky, sky_size, cube_storage), cube_texture(irradiance_function(sky), irradiance_size, cube_storage)]
    cam_pos = np.array([0.0, 2.4, -6.5], np.float32)
    return SceneSpec(
        name=f"mesh_pbr_s{seed}_{width}x{height}" + ("_texels" if texel_storage else ""), width=width, height=height,
        camera_euler=look_at_euler(cam_pos, (0.0, 1.0, 0.4)), camera_position=cam_pos, focal_length=FOV60_FOCAL,
        max_depth=max_depth, render_mode=render_mode, skybox_color=1, skybox_radiance=1, skybox_irradiance=2,
        geo_f=_stack([g[0] for g in geos]), geo_u=_stack([g[1] for g in geos]),
        mat_f=_stack(mats_f), mat_u=_stack(mats_u),
        light_f=np.zeros((0, 7), np.float32), light_u=np.zeros(0, np.uint32),
        meshes=[mesh], textures=tex)


# --- scene 3: the mesh example's lighting (Phong, Blinn, point and directional lights, meshes casting shadows) --
def lit_mesh_scene(seed: int = 0, width: int = 640, height: int = 360, max_depth: int = 5, level: int = 4,
                   texture_size: int = 128, sky_size: int = 32, irradiance_size: int = 8, leaf_size: int = 4,
                   render_mode: int = MODE_BEAUTY, ibl: bool = True) -> SceneSpec:
    """The setup of the reference's mesh example (src/examples/07_Meshes.cpp:40-110) with its missing assets
    replaced by seeded ones: a textured floor quad, a mirror sphere, three instances of a displaced icosphere
    (a Phong one and a Blinn one with 07_Meshes' phong_material / blinn_material values, both casting shadows,
    and a glass one that casts none, as 07_Meshes clears GEOMETRY_IS_SHADOWING for glass), a Phong box, a
    Lambert tetrahedron; three coloured point lights as in 07_Meshes (rim, two glass lights; all shadowing), a
    fourth point light that casts no shadows and one shadowing directional light.  `ibl` adds cube-map sky
    lighting as 07_Meshes' Cathedral skybox does.

    Draw order: per mesh instance (x, z offsets, euler y, scale); box (half-extents xyz, euler xyz); tetrahedron
    (scale, euler xyz); per point light (position jitter xyz); directional light (elevation, azimuth).
    Geometry order: floor, sphere, mesh Phong, mesh Blinn, mesh glass, box, tetrahedron.
    Textures: 0 = checker+noise albedo; with `ibl`, 1 = sky radiance cube, 2 = irradiance cube."""
    rng = PCG32(seed)
    mesh = displaced_sphere_mesh(level, 0.15, seed, leaf_size)
    geos, mats_f, mats_u = [], [], []
    # materials: 07_Meshes' Floor, Mirror, Phong, Blinn, Glass, plus a Phong box and a rough tetrahedron
    f, u = material((0.8, 0.8, 0.8), 0.2, flags=MAT_ALBEDO_MAP, texture_ids=(0,), brdf=BRDF_COOK_TORRANCE,
                    uv_repeat=(4, 4))
    mats_f.append(f); mats_u.append(u)
    f, u = material((0.0, 0.0, 0.0), 0.02, flags=MAT_REFLECTIVE, brdf=BRDF_COOK_TORRANCE,
                    reflectivity=(0.91, 0.92, 0.92), ior=float(IOR_AIR))
    mats_f.append(f); mats_u.append(u)
    f, u = material((1.0, 1.0, 1.0), 0.5, brdf=BRDF_PHONG, reflectivity=(1.0, 1.0, 0.4), ior=float(IOR_AIR))
    mats_f.append(f); mats_u.append(u)
    f, u = material((1.0, 0.3, 1.0), 0.5, brdf=BRDF_BLINN, reflectivity=(1.0, 0.4, 1.0), ior=float(IOR_AIR))
    mats_f.append(f); mats_u.append(u)
    f, u = material((0.05, 0.05, 0.05), 0.25, flags=MAT_REFRACTIVE, brdf=BRDF_COOK_TORRANCE, ior=1.52,
                    reflectivity=(0.03, 0.03, 0.03))
    mats_f.append(f); mats_u.append(u)
    f, u = material((0.4, 0.8, 0.9), 0.3, brdf=BRDF_PHONG, reflectivity=(0.9, 0.9, 0.9), ior=float(IOR_AIR))
    mats_f.append(f); mats_u.append(u)
    f, u = material((0.8, 0.8, 0.8), 0.7, brdf=BRDF_LAMBERT)
    mats_f.append(f); mats_u.append(u)

    geos.append(_geo(GEO_QUAD, (0, 0, 0), (0, 0, 0), (12, 1, 12), 0))
    geos.append(_geo(GEO_SPHERE, (0, 0, 0), (4, 1, -4), (1, 1, 1), 1))
    slots = [((-0.5, 2.4), 2, GEO_VISIBLE | GEO_SHADOWING),     # Phong
             ((3.0, 1.5), 3, GEO_VISIBLE | GEO_SHADOWING),      # Blinn
             ((-3.4, -1.6), 4, GEO_VISIBLE)]                    # glass: no shadows (07_Meshes.cpp:125-129)
    for (sx, sz), mat_id, fl in slots:
        px = sx + rng.uniform(-0.4, 0.4)
        pz = sz + rng.uniform(-0.4, 0.4)
        yaw = rng.uniform(0, 2 * math.pi)
        s = rng.uniform(1.0, 1.6)
        geos.append(_geo(GEO_MESH, (0, yaw, 0), (px, s * 1.15, pz), (s, s, s), mat_id, 0, fl))
    h = (rng.uniform(0.4, 1.0), rng.uniform(0.4, 1.0), rng.uniform(0.4, 1.0))
    eul = (rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi))
    geos.append(_geo(GEO_BOX, eul, (1.8, 1.6, -2.6), h, 5))
    s = rng.uniform(0.8, 1.4)
    eul = (rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi))
    geos.append(_geo(GEO_TET, eul, (-1.6, 1.3, -3.4), (s, s, s), 6))

    # 07_Meshes' lights (rim, glass 1, glass 2), positions jittered, plus an unshadowed fill light
    base = [((1.0, 0.25, 0.25), (6.0, 5.0, 2.0), 36.0, LIGHT_SHADOWING),
            ((0.25, 1.0, 0.25), (-3.3, 0.6, -3.0) , 20.0, LIGHT_SHADOWING),
            ((0.25, 0.25, 1.0), (-0.6, 1.75, -3.15), 20.0, LIGHT_SHADOWING),
            ((1.0, 0.9, 0.7), (-5.0, 6.0, -6.0), 25.0, 0)]
    lights_f, lights_u = [], []
    for col, (lx, ly, lz), inten, fl in base:
        p = (lx + rng.uniform(-0.3, 0.3), ly + rng.uniform(0.0, 0.3), lz + rng.uniform(-0.3, 0.3))
        lights_f.append(np.array([*col, *p, inten], np.float32))
        lights_u.append(fl)
    elev = rng.uniform(0.6, 1.1)
    azim = rng.uniform(0, 2 * math.pi)
    d = np.array([math.cos(elev) * math.cos(azim), -math.sin(elev), math.cos(elev) * math.sin(azim)], np.float32)
    d = (d / np.float32(np.linalg.norm(d.astype(np.float64)))).astype(np.float32)
    lights_f.append(np.array([0.9, 0.95, 1.0, d[0], d[1], d[2], 0.6], np.float32))   # direction of travel
    lights_u.append(LIGHT_DIRECTIONAL | LIGHT_SHADOWING)

    tex = [texture_from_texels(checker_noise_texels(texture_size, 8, seed + 202), True, True)]
    sky_ids = (-1, -1, -1)
    if ibl:
        sky = sky_function(seed + 101)
        tex += [cube_texture(sky, sky_size), cube_texture(irradiance_function(sky), irradiance_size)]
        sky_ids = (1, 1, 2)
    cam_pos = np.array([0.0, 4.5, -8.5], np.float32)
    return SceneSpec(
        name=f"lit_mesh_s{seed}_{width}x{height}", width=width, height=height,
        camera_euler=look_at_euler(cam_pos, (0.0, 0.6, 0.2)), camera_position=cam_pos,
        focal_length=FOV60_FOCAL, max_depth=max_depth, render_mode=render_mode,
        skybox_color=sky_ids[0], skybox_radiance=sky_ids[1], skybox_irradiance=sky_ids[2],
        geo_f=_stack([g[0] for g in geos]), geo_u=_stack([g[1] for g in geos]),
        mat_f=_stack(mats_f), mat_u=_stack(mats_u),
        light_f=_stack(lights_f), light_u=np.array(lights_u, np.uint32),
        meshes=[mesh], textures=tex)


# ---------------------------------------------------------------- batched views (many cameras, moving instances)
VIEW_RESOLUTIONS = ((128, 128), (320, 180), (640, 360), (1280, 720), (1920, 1080))


def instance_transforms(spec: SceneSpec, rng: PCG32) -> np.ndarray:
    """A new transform for every mesh instance of `spec` (the floor and other primitives keep theirs): Euler
    angles U[0, 2pi)^3, scale U[0.7, 1.5], position x, z U[-3.5, 3.5] resting on the floor (y = 1.15 s).
    Draw order per instance: scale, x, z, euler xyz."""
    geo_f = spec.geo_f.copy()
    for i in range(len(geo_f)):
        if spec.geo_u[i][0] != GEO_MESH:
            continue
        s = rng.uniform(0.7, 1.5)
        px, pz = rng.uniform(-3.5, 3.5), rng.uniform(-3.5, 3.5)
        eul = (rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi))
        geo_f[i, 0:3] = eul
        geo_f[i, 3:6] = (px, s * 1.15, pz)
        geo_f[i, 6:9] = (s, s, s)
    return geo_f.astype(np.float32)


def primitive_transforms(spec: SceneSpec, rng: PCG32, indices=(2, 3, 4)) -> np.ndarray:
    """New transforms for the implicit primitives `indices` of `spec` (the plain scene's spheres 1-3 by default;
    the others keep theirs): position U([-4, 4] x [0.5, 3] x [-4, 4]) and Euler angles U[0, 2pi)^3, as the plain
    scene places them; scales are kept.  Draw order per primitive: position xyz, euler xyz."""
    geo_f = spec.geo_f.copy()
    for i in indices:
        geo_f[i, 3:6] = (rng.uniform(-4, 4), rng.uniform(0.5, 3), rng.uniform(-4, 4))
        geo_f[i, 0:3] = (rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi))
    return geo_f.astype(np.float32)


def batched_views(count: int = 48, seed: int = 0, resolutions=VIEW_RESOLUTIONS):
    """`count` cameras on spheres of radius U[6, 12] around the origin (uniform direction, upper hemisphere
    kept above the floor), looking at the origin with a vertical FOV U[40, 90] degrees, at resolutions drawn
    uniformly from `resolutions`.  Returns [(width, height, euler, position, focal_length)].
    Draw order per view: resolution index, radius, direction (z = U[0.05, 1] height, azimuth U[0, 2pi)), FOV."""
    rng = PCG32(seed)
    out = []
    for _ in range(count):
        w, h = resolutions[min(int(rng.uniform() * len(resolutions)), len(resolutions) - 1)]
        r = rng.uniform(6.0, 12.0)
        up = rng.uniform(0.05, 1.0)
        az = rng.uniform(0.0, 2 * math.pi)
        horiz = math.sqrt(max(0.0, 1.0 - up * up))
        pos = np.array([r * horiz * math.cos(az), r * up + 0.5, r * horiz * math.sin(az)], np.float32)
        fov = rng.uniform(40.0, 90.0)
        focal = np.float32(1.0 / math.tan(math.radians(fov) / 2))
        out.append((w, h, look_at_euler(pos, (0.0, 0.5, 0.0)), pos, focal))
    return out


# ---------------------------------------------------------------- the large combined workload
LARGE_VIEW_RESOLUTIONS = ((3840, 2160), (2560, 1440), (1920, 1080), (1600, 900), (1280, 720), (960, 540),
                          (800, 450), (640, 360))


def normal_map_texels(size: int, seed: int, strength: float = 2.0) -> np.ndarray:
    """Tangent-space normal map of a 4-octave value-noise height field, in the channel order the reference's
    sampler reads (rotate_normal, src/slim/scene/material.h:83-86: (r, b, g) * 2 - 1 with b the up axis)."""
    y, x = np.mgrid[0:size, 0:size]
    p = np.stack([x.ravel() / size * 8.0, y.ravel() / size * 8.0, np.full(size * size, 0.5)], 1)
    h = value_noise3(p, seed, octaves=4, base_frequency=1.0).reshape(size, size)
    du = (np.roll(h, -1, 1) - np.roll(h, 1, 1)) * (size / 16.0) * strength / size * 8
    dv = (np.roll(h, -1, 0) - np.roll(h, 1, 0)) * (size / 16.0) * strength / size * 8
    n = np.stack([-du, np.ones_like(h), -dv], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    rgb = np.stack([n[..., 0], n[..., 2], n[..., 1]], -1) * 0.5 + 0.5   # r = x, g = z, b = up
    return (np.round(np.clip(rgb, 0, 1) * 255) / 255).astype(np.float32)


def large_scene(seed: int = 0, width: int = 3840, height: int = 2160, max_depth: int = 8, instances: int = 16,
                implicit: int = 32, emissive: int = 4, uv_sphere=(257, 512), texture_size: int = 2048,
                textures: int = 8, sky_size: int = 512, irradiance_size: int = 32, leaf_size: int = 4,
                render_mode: int = MODE_BEAUTY) -> SceneSpec:
    """The large combined workload: `instances` instances of one displaced UV sphere (257 rings x 512 segments =
    262,144 triangles: pole fans of 512 plus 255 rings of 1,024), `implicit` implicit primitives, `emissive`
    2x2 emissive quads (intensity U[10, 40]) facing down from y = 7..9, GGX materials with `textures` distinct
    mip-mapped albedo textures and as many normal maps, cube-map IBL, bounce depth 8.

    Instances sit on a jittered 4 x 4 grid (x, z in [-9, 9]) with Euler U[0, 2pi)^3 and scale U[0.7, 1.5];
    instance i uses texture pair i mod `textures`; instances 0 and 1 are glass (IOR 1.5) and a mirror.
    Implicit primitives (sphere, box, tetrahedron in turn) are placed at x, z U[-11, 11], y U[0.5, 4] with
    scale U[0.3, 0.9]; every fourth is a mirror, the rest Lambert or GGX.  Emissive quads are area lights
    only, shaded by the reference's heuristic estimator (shadeFromEmissiveQuads: no shadow rays).
    The albedo textures share one noise field (checker colours differ per texture), and so do the normal maps
    (each at its own strength), which keeps generation cheap at 2048^2.
    Textures: 0 .. T-1 albedo, T .. 2T-1 normal maps, 2T sky radiance cube, 2T+1 irradiance cube."""
    rng = PCG32(seed)
    rings, segs = uv_sphere
    mesh = displaced_sphere_mesh(0, 0.15, seed, leaf_size, uv_sphere=(rings, segs))
    two_pi = 2.0 * math.pi
    geos, mats_f, mats_u = [], [], []
    geos.append(_geo(GEO_QUAD, (0, 0, 0), (0, 0, 0), (12, 1, 12), 0))
    f, u = material((1, 1, 1), 0.6, flags=MAT_ALBEDO_MAP, texture_ids=(0,), brdf=BRDF_COOK_TORRANCE, uv_repeat=(6, 6))
    mats_f.append(f); mats_u.append(u)
    for i in range(instances):
        gx, gz = i % 4, i // 4
        s = rng.uniform(0.7, 1.5)
        px = -6.75 + 4.5 * gx + rng.uniform(-0.8, 0.8)
        pz = -6.75 + 4.5 * gz + rng.uniform(-0.8, 0.8)
        eul = (rng.uniform(0, two_pi), rng.uniform(0, two_pi), rng.uniform(0, two_pi))
        rough = rng.uniform(0.05, 0.9)
        metal = 1.0 if rng.uniform() < 0.5 else 0.0
        fl = GEO_VISIBLE | GEO_SHADOWING
        if i == 0:
            f, u = material((0.05, 0.05, 0.05), rough * 0.2, flags=MAT_REFRACTIVE, ior=1.5, reflectivity=(0.04, 0.04, 0.04))
            fl = GEO_VISIBLE
        elif i == 1:
            f, u = material((0.0, 0.0, 0.0), 0.02, flags=MAT_REFLECTIVE, reflectivity=(0.91, 0.92, 0.92), metalness=1.0)
        else:
            t = i % textures
            base = np.array([0.9, 0.8, 0.6], np.float32)
            f0 = 0.04 * (1 - metal) + base * metal
            f, u = material(base, rough, flags=MAT_ALBEDO_MAP | MAT_NORMAL_MAP, texture_ids=(t, textures + t),
                            reflectivity=tuple(f0), metalness=metal, normal_magnitude=0.5)
        mats_f.append(f); mats_u.append(u)
        geos.append(_geo(GEO_MESH, eul, (px, s * 1.15, pz), (s, s, s), len(geos), 0, fl))
    kinds = (GEO_SPHERE, GEO_BOX, GEO_TET)
    for j in range(implicit):
        s = rng.uniform(0.3, 0.9)
        pos = (rng.uniform(-11, 11), rng.uniform(0.5, 4.0), rng.uniform(-11, 11))
        eul = (rng.uniform(0, two_pi), rng.uniform(0, two_pi), rng.uniform(0, two_pi))
        albedo = (rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9))
        rough = rng.uniform(0.1, 0.9)
        coin = rng.uniform()
        if j % 4 == 3:
            f, u = material(albedo, 0.05, flags=MAT_REFLECTIVE, brdf=BRDF_BLINN, reflectivity=(0.9, 0.9, 0.9))
        elif coin < 0.5:
            f, u = material(albedo, 1.0, brdf=BRDF_LAMBERT)
        else:
            f, u = material(albedo, rough, brdf=BRDF_COOK_TORRANCE, reflectivity=(0.04, 0.04, 0.04))
        mats_f.append(f); mats_u.append(u)
        geos.append(_geo(kinds[j % 3], eul, pos, (s, s, s), len(geos)))
    for k in range(emissive):
        intensity = rng.uniform(10.0, 40.0)
        pos = (rng.uniform(-8, 8), rng.uniform(7, 9), rng.uniform(-8, 8))
        f, u = material((1, 1, 1), 1.0, flags=MAT_EMISSIVE, brdf=BRDF_LAMBERT, emission=(intensity,) * 3)
        mats_f.append(f); mats_u.append(u)
        geos.append(_geo(GEO_QUAD, (math.pi, 0, 0), pos, (1, 1, 1), len(geos)))   # flipped: faces down
    # textures
    n_tex = textures
    y, x = np.mgrid[0:texture_size, 0:texture_size]
    p = np.stack([x.ravel() / texture_size * 16.0, y.ravel() / texture_size * 16.0, np.zeros(texture_size ** 2)], 1)
    noise = value_noise3(p, seed + 11, octaves=4, base_frequency=1.0).reshape(texture_size, texture_size)
    tex = []
    for t in range(n_tex):
        c0 = np.array([rng.uniform(0.55, 0.95) for _ in range(3)])
        c1 = np.array([rng.uniform(0.05, 0.45) for _ in range(3)])
        squares = 4 + 2 * t
        check = ((x * squares // texture_size) + (y * squares // texture_size)) % 2
        base = np.where(check[..., None] == 1, c0, c1)
        rgb = np.clip(base * (0.8 + 0.2 * np.roll(noise, 97 * t, axis=1)[..., None]), 0, 1)
        texels = (np.round(rgb * 255) / 255).astype(np.float32)
        tex.append(texture_from_texels(texels, True, True))
    nm = normal_map_texels(texture_size, seed + 303)
    for t in range(n_tex):
        tex.append(texture_from_texels(np.roll(nm, 131 * t, axis=0), True, True))
    sky = sky_function(seed + 101)
    tex += [cube_texture(sky, sky_size), cube_texture(irradiance_function(sky), irradiance_size)]
    cam_pos = np.array([0.0, 6.0, -16.0], np.float32)
    return SceneSpec(
        name=f"large_s{seed}_{width}x{height}", width=width, height=height,
        camera_euler=look_at_euler(cam_pos, (0.0, 1.0, 0.0)), camera_position=cam_pos, focal_length=FOV60_FOCAL,
        max_depth=max_depth, render_mode=render_mode, skybox_color=2 * n_tex, skybox_radiance=2 * n_tex,
        skybox_irradiance=2 * n_tex + 1,
        geo_f=_stack([g[0] for g in geos]), geo_u=_stack([g[1] for g in geos]),
        mat_f=_stack(mats_f), mat_u=_stack(mats_u),
        light_f=np.zeros((0, 7), np.float32), light_u=np.zeros(0, np.uint32),
        meshes=[mesh], textures=tex)


def large_views(seed: int = 0, resolutions=LARGE_VIEW_RESOLUTIONS):
    """One camera per resolution, on a circle of radius U[14, 20] at height U[4, 9] looking at (0, 1, 0).
    Returns [(width, height, euler, position, focal_length)] as batched_views does."""
    rng = PCG32(seed)
    out = []
    for w, h in resolutions:
        r = rng.uniform(14.0, 20.0)
        a = rng.uniform(0.0, 2 * math.pi)
        pos = np.array([r * math.cos(a), rng.uniform(4.0, 9.0), r * math.sin(a)], np.float32)
        out.append((w, h, look_at_euler(pos, (0.0, 1.0, 0.0)), pos, FOV60_FOCAL))
    return out
