"""ctypes bindings for libslimrt.so (include/slimrt.h) and a small host-side driver.

`build_scene(spec)` derives what the reference derives on the host (orientations, mesh BVHs and triangle records,
world AABBs, the scene BVH, the camera projection) using the library's own host entry points, and `Renderer`
uploads the result and renders through the C ABI - on the GPU by default; a CPU context is the CPU side of the
reference's CPU/GPU toggle.  There is no fallback: a missing library or a missing GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libslimrt.so")

u32, i32, f32, vp = C.c_uint32, C.c_int32, C.c_float, C.c_void_p

SLIMRT_DEVICE_GPU, SLIMRT_DEVICE_CPU = 0, 1
FLAG_STACK_OVERFLOW, FLAG_QUEUE_OVERFLOW, FLAG_NAN_SAMPLE = 1, 2, 4   # slimrt_frame_flags bits
OUTPUT_CANVAS, OUTPUT_RADIANCE, OUTPUT_HITS = 0, 1, 2
RAY_CONES = 1
ERRORS = {-1: "SLIMRT_E_ARG", -2: "SLIMRT_E_DEVICE", -3: "SLIMRT_E_MEMORY", -4: "SLIMRT_E_STATE", -5: "SLIMRT_E_OVERFLOW"}

# numpy mirrors of the C structs
GEOMETRY = np.dtype([("orientation", "<f4", 4), ("position", "<f4", 3), ("scale", "<f4", 3), ("type", "<u4"),
                     ("material_id", "<u4"), ("id", "<u4"), ("flags", "<u4")])
MATERIAL = np.dtype([("albedo", "<f4", 3), ("roughness", "<f4"), ("emission", "<f4", 3), ("ior", "<f4"),
                     ("reflectivity", "<f4", 3), ("metalness", "<f4"), ("uv_repeat", "<f4", 2),
                     ("normal_magnitude", "<f4"), ("flags", "<u4"), ("texture_count", "<u4"),
                     ("texture_ids", "<u4", 2), ("brdf", "<u4")])
LIGHT = np.dtype([("color", "<f4", 3), ("intensity", "<f4"), ("position", "<f4", 3), ("flags", "<u4")])
BVH_NODE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3), ("first_index", "<u4"), ("leaf_count", "<u2"),
                     ("depth", "u1"), ("flags", "u1")])
TRIANGLE = np.dtype([("local_to_tangent", "<f4", 9), ("position", "<f4", 3), ("normal", "<f4", 3),
                     ("n1", "<f4", 3), ("n2", "<f4", 3), ("n3", "<f4", 3), ("uv1", "<f4", 2), ("uv2", "<f4", 2),
                     ("uv3", "<f4", 2), ("uv_coverage", "<f4"), ("padding", "<f4")])
assert GEOMETRY.itemsize == 56 and MATERIAL.itemsize == 80 and LIGHT.itemsize == 32
BVH4_NODE = np.dtype([("min", "<f4", (3, 4)), ("max", "<f4", (3, 4)), ("child", "<u4", 4), ("count", "<u4", 4)])
assert BVH_NODE.itemsize == 32 and TRIANGLE.itemsize == 128 and BVH4_NODE.itemsize == 128
OVERLAY_TRANSFORM = np.dtype([("orientation", "<f4", 4), ("position", "<f4", 3), ("scale", "<f4", 3)])
OVERLAY_BOX = np.dtype([("transform", OVERLAY_TRANSFORM), ("min", "<f4", 3), ("max", "<f4", 3), ("color", "<f4", 3),
                        ("opacity", "<f4"), ("line_width", "<u4"), ("sides", "<u4")])
assert OVERLAY_BOX.itemsize == 88
BOX_ALL_SIDES = 63
BOX_LEFT, BOX_BOTTOM, BOX_BACK, BOX_RIGHT, BOX_TOP, BOX_FRONT = 1, 2, 4, 8, 16, 32


class Mesh(C.Structure):
    _fields_ = [("aabb_min", f32 * 3), ("aabb_max", f32 * 3), ("nodes", vp), ("node_count", u32), ("height", u32),
                ("triangles", vp), ("triangle_count", u32), ("normals_count", u32), ("uvs_count", u32)]


class TextureMip(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("texel_quads", vp)]


class Texture(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("flags", u32), ("mip_count", u32), ("mips", vp)]


class SceneDesc(C.Structure):
    _fields_ = [("geometries", vp), ("geometry_count", u32), ("materials", vp), ("material_count", u32),
                ("lights", vp), ("light_count", u32), ("meshes", vp), ("mesh_count", u32),
                ("textures", vp), ("texture_count", u32), ("scene_nodes", vp), ("scene_node_count", u32),
                ("scene_height", u32), ("scene_leaf_ids", vp)]


class RenderParams(C.Structure):
    _fields_ = [("inverted_camera_rotation", f32 * 9), ("start", f32 * 3), ("right", f32 * 3), ("down", f32 * 3),
                ("camera_position", f32 * 3), ("C_start", f32 * 2), ("squared_distance_to_projection_plane", f32),
                ("sample_size", f32), ("width", u32), ("height", u32), ("ssaa", u32), ("max_depth", u32),
                ("render_mode", u32), ("skybox_color", i32), ("skybox_radiance", i32), ("skybox_irradiance", i32),
                ("output", u32), ("flags", u32), ("mip_level_colors", (f32 * 3) * 9)]


class PickResult(C.Structure):
    _fields_ = [("geometry", i32), ("light", i32), ("triangle", i32), ("distance", f32), ("position", f32 * 3)]


class OverlayView(C.Structure):
    _fields_ = [("camera_euler", f32 * 3), ("camera_position", f32 * 3), ("focal_length", f32), ("width", u32),
                ("height", u32), ("ssaa", u32), ("near_distance", f32), ("far_distance", f32)]


def overlay_view(spec: scenes.SceneSpec, near: float = 0.001, far: float = 1000.0) -> OverlayView:
    """The viewport of a scene's canvas with the reference's default clipping distances."""
    return OverlayView((f32 * 3)(*np.asarray(spec.camera_euler, np.float32)),
                       (f32 * 3)(*np.asarray(spec.camera_position, np.float32)), float(spec.focal_length),
                       spec.width, spec.height, int(spec.ssaa), near, far)


class ViewsConfig(C.Structure):
    _fields_ = [("tile_shape", u32), ("blocks_per_cu", u32), ("tiles_per_grab", u32), ("persistent", u32),
                ("node_width", u32)]


SCENE_REFIT, SCENE_REBUILD = 0, 1
WAVEFRONT, WAVEFRONT_PER_BOUNCE, WAVEFRONT_SORT, BVH4, WAVEFRONT_COUNT, WAVEFRONT_SPLIT = 2, 4, 8, 64, 128, 256   # flags
TILE_ORDER_FIXED = 512   # per-sample frames: tiles in image order instead of costliest first
BVH8 = 1024   # per-sample frames and the CPU context: 8-wide mesh BVHs with quantised child boxes


def WAVEFRONT_BANDS(log2n: int) -> int:
    return (log2n & 3) << 4


# RayTracingRenderer's mip_level_colors (src/slim/renderer/renderer.h:50-58) as Color(ColorID) values
MIP_LEVEL_COLORS = [(1.0, 0.5, 0.5), (0.75, 0.75, 0.0), (0.5, 1.0, 0.5), (0.75, 0.0, 0.75), (0.0, 0.75, 0.75),
                    (0.5, 0.5, 1.0), (0.75, 0.75, 0.75), (0.5, 0.5, 0.5), (0.25, 0.25, 0.25)]

_lib = None


def lib():
    """Load libslimrt.so; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "or make -C slimtracincpp_amd/csrc")
        # torch's ROCm wheel carries its own HIP runtime (soname libamdhip64.so.7, loaded by file name).  Loaded
        # first, it is the one the library binds to as well; loaded after the library, the process would hold two
        # HIP runtimes and the second to start sees no device.  So torch, when installed, is imported first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        sig = {
            "slimrt_create": [P(vp), u32], "slimrt_destroy": [vp], "slimrt_set_stream": [vp, vp],
            "slimrt_upload_scene": [vp, P(SceneDesc)], "slimrt_upload_geometries": [vp, vp, u32],
            "slimrt_upload_materials": [vp, vp, u32], "slimrt_upload_lights": [vp, vp, u32],
            "slimrt_upload_scene_bvh": [vp, vp, u32, vp, u32],
            "slimrt_render": [vp, P(RenderParams), vp, vp], "slimrt_render_device": [vp, P(RenderParams), vp, vp],
            "slimrt_render_device_noalloc": [vp, P(RenderParams), vp, vp],
            "slimrt_synchronize": [vp], "slimrt_count_work": [vp, P(RenderParams), vp],
            "slimrt_frame_flags": [vp, P(u32)], "slimrt_set_memory_limit": [vp, C.c_uint64],
            "slimrt_memory_usage": [vp, P(C.c_uint64), P(C.c_uint64)],
            "slimrt_last_kernel_ms": [vp, P(f32)],
            "slimrt_host_orientation": [vp, vp],
            "slimrt_host_projection": [vp, vp, f32, u32, u32, u32, P(RenderParams)],
            "slimrt_host_world_aabbs": [vp, u32, vp, u32, vp],
            "slimrt_host_build_bvh": [vp, u32, u32, vp, P(u32), P(u32), vp],
            "slimrt_host_build_mesh": [vp, u32, vp, vp, u32, vp, vp, u32, vp, u32, u32, vp, P(u32), P(u32), vp, vp],
            "slimrt_update_scene_device": [vp, vp, u32, u32],
            "slimrt_download_scene_bvh": [vp, vp, u32, vp, u32, P(u32)],
            "slimrt_views_create": [vp, vp, u32, vp, vp, P(vp)], "slimrt_views_destroy": [vp],
            "slimrt_views_configure": [vp, P(ViewsConfig)], "slimrt_views_get_config": [vp, P(ViewsConfig)],
            "slimrt_views_frame": [vp, vp, u32, u32, u32], "slimrt_views_last_ms": [vp, P(f32)],
            "slimrt_views_autotune": [vp, u32, P(ViewsConfig), P(f32)],
            "slimrt_render_autotune": [vp, P(RenderParams), vp, vp, u32, P(u32), P(f32)],
            "slimrt_host_collapse_bvh4": [vp, u32, u32, vp, u32, P(u32), P(u32), P(u32)],
            "slimrt_host_collapse_bvh8": [vp, u32, vp, u32, vp, u32, P(u32), vp, u32, P(u32), P(u32)],
            "slimrt_wavefront_counts": [vp, vp, u32],
            "slimrt_tile_costs": [vp, vp, vp, u32, P(u32), P(u32)],
            "slimrt_pick": [vp, P(RenderParams), i32, i32, P(PickResult)],
            "slimrt_canvas_content": [vp, vp, u32, u32, u32, vp],
            "slimrt_host_bake_texture": [vp, u32, u32, u32, vp, C.c_size_t, P(C.c_size_t), P(u32), vp, P(u32)],
            "slimrt_host_texels_from_quads": [vp, u32, u32, u32, vp, P(C.c_uint64)],
            "slimrt_draw_boxes": [vp, P(OverlayView), vp, u32, vp, vp],
            "slimrt_draw_mesh_wireframe": [vp, P(OverlayView), vp, u32, vp, u32, vp, vp, f32, u32, vp, vp],
            "slimrt_host_bvh_boxes": [vp, u32, vp, u32, u32, f32, u32, vp, P(u32)],
            "slimrt_host_camera_rotation": [vp, vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.slimrt_host_float_to_half.argtypes = [f32]
        L.slimrt_host_float_to_half.restype = C.c_uint16
        L.slimrt_last_error.argtypes = [vp]
        L.slimrt_last_error.restype = C.c_char_p
        L.slimrt_views_last_error.argtypes = [vp]
        L.slimrt_views_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class SlimrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def host_texels_from_quads(q: np.ndarray, width: int, height: int, texel_storage: int):
    """slimrt_host_texels_from_quads: (bordered texel grid (h+2, w+2, 4) u8 or float16, mismatching quads)."""
    q = np.ascontiguousarray(q, np.uint8)
    assert q.size == 12 * (width + 1) * (height + 1)
    out = np.zeros((height + 2, width + 2, 4), np.uint8 if texel_storage == scenes.TEXELS_RGBA8 else np.float16)
    bad = C.c_uint64(0)
    _check(lib().slimrt_host_texels_from_quads(_ptr(q), width, height, texel_storage, _ptr(out), C.byref(bad)),
           None, "slimrt_host_texels_from_quads")
    return out, int(bad.value)


def host_float_to_half(x: float) -> int:
    """slimrt_host_float_to_half: the IEEE half bits nearest to the float x (ties to even)."""
    return int(lib().slimrt_host_float_to_half(float(x)))


def mip_bytes(q: np.ndarray) -> np.ndarray:
    """A mip's storage as contiguous bytes: u8 quads or RGBA8 texels as they are, float16 texels reinterpreted."""
    q = np.ascontiguousarray(q)
    if q.dtype == np.float16:
        return q.view(np.uint8)
    assert q.dtype == np.uint8, q.dtype
    return q


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


def _check(rc, ctx=None, what=""):
    if rc != 0:
        msg = lib().slimrt_last_error(ctx).decode() if ctx else what
        raise SlimrtError(rc, msg or what)


# --- host-side scene derivation (through the library's host entry points) -------------------------------------
def orientation(euler) -> np.ndarray:
    e = np.ascontiguousarray(euler, np.float32)
    q = np.zeros(4, np.float32)
    _check(lib().slimrt_host_orientation(_ptr(e), _ptr(q)), what="orientation")
    return q


def camera_rotation(euler) -> np.ndarray:
    """The camera's rotation matrix (rows X, Y, Z) for Euler angles (slimrt_host_camera_rotation)."""
    e = np.ascontiguousarray(euler, np.float32)
    r = np.zeros(9, np.float32)
    _check(lib().slimrt_host_camera_rotation(_ptr(e), _ptr(r)), what="camera_rotation")
    return r.reshape(3, 3)


def overlay_transform(orientation=(0, 0, 0, 1), position=(0, 0, 0), scale=(1, 1, 1)) -> np.ndarray:
    t = np.zeros((), OVERLAY_TRANSFORM)
    t["orientation"], t["position"], t["scale"] = orientation, position, scale
    return t


def bvh_boxes(nodes: np.ndarray, transform: np.ndarray, min_depth: int = 0, max_depth: int = 5,
              opacity: float = 0.25, line_width: int = 1) -> np.ndarray:
    """drawBVH's boxes (slimrt_host_bvh_boxes) for slimrt_draw_boxes."""
    a = np.ascontiguousarray(nodes, BVH_NODE)
    t = np.ascontiguousarray(transform, OVERLAY_TRANSFORM)
    out = np.zeros(max(len(a), 1), OVERLAY_BOX)
    n = u32()
    _check(lib().slimrt_host_bvh_boxes(_ptr(a), len(a), _ptr(t), min_depth, max_depth, opacity, line_width, _ptr(out),
                                       C.byref(n)), what="bvh_boxes")
    return out[:n.value]


def build_bvh(aabbs: np.ndarray, max_leaf: int = 1):
    a = np.ascontiguousarray(aabbs, np.float32).reshape(-1, 6)
    n = a.shape[0]
    nodes = np.zeros(max(2 * n - 1, 1), BVH_NODE)
    leaf = np.zeros(n, np.uint32)
    cnt, h = u32(), u32()
    _check(lib().slimrt_host_build_bvh(_ptr(a), n, max_leaf, _ptr(nodes), C.byref(cnt), C.byref(h), _ptr(leaf)),
           what="build_bvh")
    return nodes[:cnt.value], leaf, h.value


def collapse_bvh4(nodes: np.ndarray, triangle_count: int, with_stack: bool = False):
    """The 4-wide tree the renderer walks for a mesh BVH under BVH4 (slimrt_host_collapse_bvh4): nodes, height
    (and the stack entries its walk can need, with_stack)."""
    a = np.ascontiguousarray(nodes, BVH_NODE)
    out = np.zeros((len(a) + 1) // 2, BVH4_NODE)
    cnt, h, st = u32(), u32(), u32()
    _check(lib().slimrt_host_collapse_bvh4(_ptr(a), len(a), triangle_count, _ptr(out), len(out), C.byref(cnt),
                                           C.byref(h), C.byref(st)), what="collapse_bvh4")
    return (out[:cnt.value], h.value, st.value) if with_stack else (out[:cnt.value], h.value)


BVH8_NODE = np.dtype([("origin", "<f4", 3), ("exps", "<u4"), ("child", "<u4", 8), ("count", "u1", 8),
                      ("qlo", "u1", (3, 8)), ("qhi", "u1", (3, 8)), ("pad", "<u4", 6)])
assert BVH8_NODE.itemsize == 128


def collapse_bvh8(nodes: np.ndarray, triangles: np.ndarray):
    """The 8-wide tree the renderer walks for a mesh BVH under BVH8 (slimrt_host_collapse_bvh8): nodes, leaf records
    (as uint32 words, 4 per 16-byte unit) and the stack entries its walk can need."""
    a = np.ascontiguousarray(nodes, BVH_NODE)
    t = np.ascontiguousarray(triangles, TRIANGLE)
    ncnt, lcnt, st = u32(), u32(), u32()
    rc = lib().slimrt_host_collapse_bvh8(_ptr(a), len(a), _ptr(t), len(t), None, 0, C.byref(ncnt), None, 0,
                                         C.byref(lcnt), C.byref(st))
    if rc not in (0, -5):
        _check(rc, what="collapse_bvh8")
    out = np.zeros(ncnt.value, BVH8_NODE)
    leaves = np.zeros(4 * lcnt.value, np.uint32)
    _check(lib().slimrt_host_collapse_bvh8(_ptr(a), len(a), _ptr(t), len(t), _ptr(out), len(out), C.byref(ncnt),
                                           _ptr(leaves), lcnt.value, C.byref(lcnt), C.byref(st)), what="collapse_bvh8")
    return out, leaves, st.value


class MeshBuild:
    def __init__(self, m: scenes.MeshSpec):
        T = m.triangle_count
        self.nodes = np.zeros(max(2 * T - 1, 1), BVH_NODE)
        self.triangles = np.zeros(T, TRIANGLE)
        self.aabb = np.zeros(6, np.float32)
        pos = np.ascontiguousarray(m.positions, np.float32)
        pidx = np.ascontiguousarray(m.position_indices, np.uint32)
        nrm = None if m.normals is None else np.ascontiguousarray(m.normals, np.float32)
        nidx = None if m.normal_indices is None else np.ascontiguousarray(m.normal_indices, np.uint32)
        uv = None if m.uvs is None else np.ascontiguousarray(m.uvs, np.float32)
        uidx = None if m.uv_indices is None else np.ascontiguousarray(m.uv_indices, np.uint32)
        cnt, h = u32(), u32()
        _check(lib().slimrt_host_build_mesh(
            _ptr(pos), pos.shape[0], _ptr(pidx), _ptr(nrm), 0 if nrm is None else nrm.shape[0], _ptr(nidx),
            _ptr(uv), 0 if uv is None else uv.shape[0], _ptr(uidx), T, m.leaf_size,
            _ptr(self.nodes), C.byref(cnt), C.byref(h), _ptr(self.triangles), _ptr(self.aabb)), what="build_mesh")
        self.nodes = self.nodes[:cnt.value]
        self.height = h.value
        self.normals_count = 0 if nrm is None else nrm.shape[0]
        self.uvs_count = 0 if uv is None else uv.shape[0]

    def desc(self) -> Mesh:
        return Mesh((f32 * 3)(*self.aabb[:3]), (f32 * 3)(*self.aabb[3:]), _ptr(self.nodes), len(self.nodes),
                    self.height, _ptr(self.triangles), len(self.triangles), self.normals_count, self.uvs_count)


def geometries_from_spec(spec: scenes.SceneSpec, geo_f: np.ndarray | None = None) -> np.ndarray:
    gf = spec.geo_f if geo_f is None else geo_f
    g = np.zeros(spec.geometry_count, GEOMETRY)
    for i in range(spec.geometry_count):
        g[i]["orientation"] = orientation(gf[i, 0:3])
    g["position"] = gf[:, 3:6]
    g["scale"] = gf[:, 6:9]
    g["type"] = spec.geo_u[:, 0]
    g["material_id"] = spec.geo_u[:, 1]
    g["id"] = spec.geo_u[:, 2]
    g["flags"] = spec.geo_u[:, 3]
    return g


def materials_from_spec(spec: scenes.SceneSpec) -> np.ndarray:
    m = np.zeros(spec.mat_f.shape[0], MATERIAL)
    f, u = spec.mat_f, spec.mat_u
    m["albedo"] = f[:, scenes.MAT_F_ALBEDO]
    m["roughness"] = f[:, scenes.MAT_F_ROUGHNESS]
    m["normal_magnitude"] = f[:, scenes.MAT_F_NORMAL_MAGNITUDE]
    m["uv_repeat"] = f[:, scenes.MAT_F_UV_REPEAT]
    m["emission"] = f[:, scenes.MAT_F_EMISSION]
    m["ior"] = f[:, scenes.MAT_F_IOR]
    m["reflectivity"] = f[:, scenes.MAT_F_REFLECTIVITY]
    m["metalness"] = f[:, scenes.MAT_F_METALNESS]
    m["flags"] = u[:, 0] & 0xFF
    m["texture_count"] = u[:, 1] & 0xFF
    m["texture_ids"] = u[:, 2:4] & 0xFF
    m["brdf"] = u[:, 4]
    return m


def lights_from_spec(spec: scenes.SceneSpec) -> np.ndarray:
    n = 0 if spec.light_f is None else spec.light_f.shape[0]
    lt = np.zeros(n, LIGHT)
    if n:
        lt["color"] = spec.light_f[:, 0:3]
        lt["position"] = spec.light_f[:, 3:6]
        lt["intensity"] = spec.light_f[:, 6]
        lt["flags"] = spec.light_u & 0xFF
    return lt


class SceneBuild:
    """Everything the host derives from a spec before upload, kept alive for the C structs that point into it."""

    def __init__(self, spec: scenes.SceneSpec):
        self.spec = spec
        self.meshes = [MeshBuild(m) for m in spec.meshes]
        self.materials = materials_from_spec(spec)
        self.lights = lights_from_spec(spec)
        self.mesh_descs = (Mesh * max(len(self.meshes), 1))(*[m.desc() for m in self.meshes])
        self.quads = [[mip_bytes(q) for q in t.mips] for t in spec.textures]
        self.mip_descs = [(TextureMip * len(t.mips))(*[TextureMip(w, h, _ptr(q)) for (w, h), q in zip(t.mip_dims, qs)])
                          for t, qs in zip(spec.textures, self.quads)]
        self.tex_descs = (Texture * max(len(spec.textures), 1))(
            *[Texture(t.width, t.height, t.flags, len(t.mips), C.cast(md, vp)) for t, md in zip(spec.textures, self.mip_descs)])
        self.set_transforms(spec.geo_f)
        self.params = self.projection(spec.camera_euler, spec.camera_position, spec.focal_length,
                                      spec.width, spec.height, spec.ssaa)

    def set_transforms(self, geo_f: np.ndarray):
        """Scene::updateAABBs + Scene::updateBVH for new transforms (scene.h:165-191)."""
        self.geometries = geometries_from_spec(self.spec, geo_f)
        self.aabbs = np.zeros((len(self.geometries), 6), np.float32)
        _check(lib().slimrt_host_world_aabbs(_ptr(self.geometries), len(self.geometries),
                                             C.cast(self.mesh_descs, vp), len(self.meshes), _ptr(self.aabbs)),
               what="world_aabbs")
        self.scene_nodes, self.leaf_ids, self.scene_height = build_bvh(self.aabbs, 1)

    def projection(self, euler, position, focal, width, height, ssaa=False, max_depth=None, render_mode=None):
        p = RenderParams()
        e = np.ascontiguousarray(euler, np.float32)
        pos = np.ascontiguousarray(position, np.float32)
        _check(lib().slimrt_host_projection(_ptr(e), _ptr(pos), float(focal), width, height, int(ssaa), C.byref(p)),
               what="projection")
        s = self.spec
        p.max_depth = s.max_depth if max_depth is None else max_depth
        p.render_mode = s.render_mode if render_mode is None else render_mode
        p.skybox_color, p.skybox_radiance, p.skybox_irradiance = s.skybox_color, s.skybox_radiance, s.skybox_irradiance
        p.output = OUTPUT_RADIANCE
        p.flags = 0
        for i, c in enumerate(MIP_LEVEL_COLORS):
            for j in range(3):
                p.mip_level_colors[i][j] = c[j]
        return p

    def desc(self) -> SceneDesc:
        s = self.spec
        return SceneDesc(_ptr(self.geometries), len(self.geometries), _ptr(self.materials), len(self.materials),
                         _ptr(self.lights), len(self.lights), C.cast(self.mesh_descs, vp), len(self.meshes),
                         C.cast(self.tex_descs, vp), len(s.textures), _ptr(self.scene_nodes), len(self.scene_nodes),
                         self.scene_height, _ptr(self.leaf_ids))


def build_scene(spec: scenes.SceneSpec) -> SceneBuild:
    return SceneBuild(spec)


class Renderer:
    """A slimrt context with an uploaded scene.  device='gpu' (default) or 'cpu'."""

    def __init__(self, scene: SceneBuild | scenes.SceneSpec, device: str = "gpu"):
        self.scene = scene if isinstance(scene, SceneBuild) else SceneBuild(scene)
        self.ctx = vp()
        flags = SLIMRT_DEVICE_CPU if device == "cpu" else SLIMRT_DEVICE_GPU
        rc = lib().slimrt_create(C.byref(self.ctx), flags)
        if rc != 0:
            raise SlimrtError(rc, f"slimrt_create({device}) failed")
        self.device = device
        d = self.scene.desc()
        _check(lib().slimrt_upload_scene(self.ctx, C.byref(d)), self.ctx)

    def close(self):
        if self.ctx:
            lib().slimrt_destroy(self.ctx)
            self.ctx = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self, mode=None, max_depth=None, output=OUTPUT_RADIANCE, ray_cones=False, camera=None, size=None):
        s = self.scene.spec
        euler, pos, focal = (s.camera_euler, s.camera_position, s.focal_length) if camera is None else camera
        w, h = (s.width, s.height) if size is None else size
        p = self.scene.projection(euler, pos, focal, w, h, s.ssaa, max_depth, mode)
        p.output = output
        p.flags = RAY_CONES if ray_cones else 0
        return p

    def render(self, params: RenderParams | None = None, **kw):
        p = params if params is not None else self.params(**kw)
        n = p.width * p.height * (4 if p.ssaa else 1)
        px = np.zeros((n, 4), np.float32)
        dp = np.zeros(n, np.float32)
        _check(lib().slimrt_render(self.ctx, C.byref(p), _ptr(px), _ptr(dp)), self.ctx)
        ws, hs = p.width * (2 if p.ssaa else 1), p.height * (2 if p.ssaa else 1)
        if p.ssaa:
            return px, dp
        return px.reshape(hs, ws, 4), dp.reshape(hs, ws)

    def primary_hits(self, **kw):
        """Primary closest hit per sample: geometry index, leaf-ordered triangle id, distance (OUTPUT_HITS)."""
        px, _ = self.render(output=OUTPUT_HITS, **kw)
        flat = px.reshape(-1, 4)
        shape = px.shape[:-1]
        return (flat[:, 0].view(np.int32).reshape(shape), flat[:, 1].view(np.int32).reshape(shape),
                flat[:, 2].copy().reshape(shape))

    def render_device(self, params: RenderParams, pixels_ptr: int, depths_ptr: int):
        _check(lib().slimrt_render_device(self.ctx, C.byref(params), C.c_void_p(pixels_ptr), C.c_void_p(depths_ptr)),
               self.ctx)

    def render_device_noalloc(self, params: RenderParams, pixels_ptr: int, depths_ptr: int):
        """slimrt_render_device without any allocation (per-sample frames; scratch in a static per-GPU array)."""
        _check(lib().slimrt_render_device_noalloc(self.ctx, C.byref(params), C.c_void_p(pixels_ptr),
                                                  C.c_void_p(depths_ptr)), self.ctx)

    def autotune_form(self, params: RenderParams, pixels_ptr: int, depths_ptr: int, reps: int = 3):
        """Fastest frame form (megakernel or wavefront variants) for these params: (flags, median ms)."""
        flags, ms = u32(), f32()
        _check(lib().slimrt_render_autotune(self.ctx, C.byref(params), C.c_void_p(pixels_ptr), C.c_void_p(depths_ptr),
                                            reps, C.byref(flags), C.byref(ms)), self.ctx)
        return flags.value, ms.value

    def synchronize(self):
        _check(lib().slimrt_synchronize(self.ctx), self.ctx)

    def frame_flags(self) -> int:
        """slimrt_frame_flags: FLAG_* bits of the frames since the last call (then cleared)."""
        f = u32()
        _check(lib().slimrt_frame_flags(self.ctx, C.byref(f)), self.ctx)
        return f.value

    def set_memory_limit(self, nbytes: int):
        _check(lib().slimrt_set_memory_limit(self.ctx, nbytes), self.ctx)

    def memory_usage(self):
        """(bytes of device memory the context holds, its limit or 0)"""
        used, limit = C.c_uint64(), C.c_uint64()
        _check(lib().slimrt_memory_usage(self.ctx, C.byref(used), C.byref(limit)), self.ctx)
        return used.value, limit.value

    def set_stream(self, stream_ptr: int | None):
        _check(lib().slimrt_set_stream(self.ctx, C.c_void_p(stream_ptr) if stream_ptr else None), self.ctx)

    def last_kernel_ms(self) -> float:
        ms = f32()
        _check(lib().slimrt_last_kernel_ms(self.ctx, C.byref(ms)), self.ctx)
        return ms.value

    def count_work(self, params: RenderParams | None = None, **kw) -> dict:
        p = params if params is not None else self.params(**kw)
        c = np.zeros(8, np.uint64)
        _check(lib().slimrt_count_work(self.ctx, C.byref(p), _ptr(c)), self.ctx)
        names = ("rays", "pairs", "roots", "instances", "triangles", "mesh_hits", "taps", "pixels")
        return dict(zip(names, (int(v) for v in c)))

    def pick(self, x: int, y: int, params: RenderParams | None = None) -> dict:
        """slimrt_pick: what Selection's click at pixel (x, y) finds: geometry, light, triangle, distance, position."""
        p = params if params is not None else self.params()
        res = PickResult()
        _check(lib().slimrt_pick(self.ctx, C.byref(p), x, y, C.byref(res)), self.ctx)
        return {"geometry": res.geometry, "light": res.light, "triangle": res.triangle,
                "distance": res.distance, "position": np.array(res.position[:], np.float32)}

    def canvas_content(self, pixels_ptr: int, width: int, height: int, ssaa: bool, content_ptr: int):
        """slimrt_canvas_content: window content words of a canvas (device pointers on a GPU context)."""
        _check(lib().slimrt_canvas_content(self.ctx, C.c_void_p(pixels_ptr), width, height, int(ssaa),
                                           C.c_void_p(content_ptr)), self.ctx)

    def draw_boxes(self, view: OverlayView, boxes: np.ndarray, pixels_ptr: int, depths_ptr: int):
        """slimrt_draw_boxes into a canvas (device pointers on a GPU context, host pointers on a CPU one)."""
        b = np.ascontiguousarray(boxes, OVERLAY_BOX)
        _check(lib().slimrt_draw_boxes(self.ctx, C.byref(view), _ptr(b), len(b), C.c_void_p(pixels_ptr),
                                       C.c_void_p(depths_ptr)), self.ctx)

    def draw_mesh_wireframe(self, view: OverlayView, positions: np.ndarray, edges: np.ndarray, transform: np.ndarray,
                            color, opacity: float, line_width: int, pixels_ptr: int, depths_ptr: int):
        """slimrt_draw_mesh_wireframe: drawMesh's edges into a canvas."""
        pos = np.ascontiguousarray(positions, np.float32)
        e = np.ascontiguousarray(edges, np.uint32)
        t = np.ascontiguousarray(transform, OVERLAY_TRANSFORM)
        col = np.ascontiguousarray(color, np.float32)
        _check(lib().slimrt_draw_mesh_wireframe(self.ctx, C.byref(view), _ptr(pos), pos.size // 3, _ptr(e),
                                                e.size // 2, _ptr(t), _ptr(col), opacity, line_width,
                                                C.c_void_p(pixels_ptr), C.c_void_p(depths_ptr)), self.ctx)

    def wavefront_counts(self) -> np.ndarray:
        """Path counters of the last frame rendered with WAVEFRONT | WAVEFRONT_COUNT: paths per integrator
        iteration (32), then bounce rays spawned, samples finished."""
        c = np.zeros(34, np.uint32)
        _check(lib().slimrt_wavefront_counts(self.ctx, _ptr(c), len(c)), self.ctx)
        return c

    def tile_costs(self):
        """Per-tile durations (wall-clock ticks) of the last per-sample frame rendered without TILE_ORDER_FIXED and
        that frame's dispatch order (order[k] = the tile dispatched k-th); empty arrays before the first such frame.
        tile_resident holds the frame kernel's resident waves the order was weighed against."""
        n, res = u32(0), u32(0)
        _check(lib().slimrt_tile_costs(self.ctx, None, None, 0, C.byref(n), C.byref(res)), self.ctx)
        costs, order = np.zeros(n.value, np.uint32), np.zeros(n.value, np.uint32)
        if n.value:
            _check(lib().slimrt_tile_costs(self.ctx, _ptr(costs), _ptr(order), n.value, C.byref(n), None), self.ctx)
        self.tile_resident = res.value
        return costs, order

    def update_scene_device(self, geo_f: np.ndarray, mode: int = SCENE_REFIT):
        """New transforms for every geometry; world AABBs and the scene BVH (refit or rebuild) on the device."""
        g = geometries_from_spec(self.scene.spec, geo_f)
        _check(lib().slimrt_update_scene_device(self.ctx, _ptr(g), len(g), mode), self.ctx)
        self._last_geometries = g   # the copy is asynchronous: keep the source alive until the next call

    def download_scene_bvh(self):
        n_cap = 2 * len(self.scene.geometries) + 1
        nodes = np.zeros(n_cap, BVH_NODE)
        leaf = np.zeros(len(self.scene.geometries), np.uint32)
        n = u32()
        _check(lib().slimrt_download_scene_bvh(self.ctx, _ptr(nodes), n_cap, _ptr(leaf), len(leaf), C.byref(n)), self.ctx)
        return nodes[:n.value], leaf

    def update_geometries(self, geo_f: np.ndarray):
        """Per-frame transform update: host AABBs + SAH rebuild, then uploadGeometries + uploadSceneBVH."""
        sc = self.scene
        sc.set_transforms(geo_f)
        _check(lib().slimrt_upload_geometries(self.ctx, _ptr(sc.geometries), len(sc.geometries)), self.ctx)
        _check(lib().slimrt_upload_scene_bvh(self.ctx, _ptr(sc.scene_nodes), len(sc.scene_nodes),
                                             _ptr(sc.leaf_ids), sc.scene_height), self.ctx)


class Views:
    """Many camera views of a Renderer's scene traced in one launch (include/slimrt_views.h).
    pixels/depths: one (device pointer) pair per view, e.g. torch tensors' data_ptr()."""

    def __init__(self, renderer: Renderer, params: list, pixels_ptrs: list, depths_ptrs: list):
        self.r = renderer
        self.params = (RenderParams * len(params))(*params)
        self._px = (vp * len(params))(*[vp(p) for p in pixels_ptrs])
        self._dp = (vp * len(params))(*[vp(p) for p in depths_ptrs])
        self.h = vp()
        _check(lib().slimrt_views_create(renderer.ctx, C.cast(self.params, vp), len(params), C.cast(self._px, vp),
                                         C.cast(self._dp, vp), C.byref(self.h)), renderer.ctx)
        self._geos = None

    def _check(self, rc):
        if rc != 0:
            raise SlimrtError(rc, lib().slimrt_views_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().slimrt_views_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, tile_shape=0, blocks_per_cu=2, tiles_per_grab=1, persistent=0, node_width=0):
        cfg = ViewsConfig(tile_shape, blocks_per_cu, tiles_per_grab, persistent, node_width)
        self._check(lib().slimrt_views_configure(self.h, C.byref(cfg)))

    def config(self) -> dict:
        cfg = ViewsConfig()
        self._check(lib().slimrt_views_get_config(self.h, C.byref(cfg)))
        return {k: getattr(cfg, k) for k, _ in ViewsConfig._fields_}

    def frame(self, geo_f: np.ndarray | None = None, mode: int = SCENE_REFIT, graph: bool = False):
        if geo_f is None:
            self._check(lib().slimrt_views_frame(self.h, None, 0, mode, int(graph)))
        else:
            self._geos = geometries_from_spec(self.r.scene.spec, geo_f)
            self._check(lib().slimrt_views_frame(self.h, _ptr(self._geos), len(self._geos), mode, int(graph)))

    def last_ms(self) -> float:
        ms = f32()
        self._check(lib().slimrt_views_last_ms(self.h, C.byref(ms)))
        return ms.value

    def autotune(self, reps: int = 3):
        cfg = ViewsConfig()
        ms = f32()
        self._check(lib().slimrt_views_autotune(self.h, reps, C.byref(cfg), C.byref(ms)))
        return {k: getattr(cfg, k) for k, _ in ViewsConfig._fields_}, ms.value
