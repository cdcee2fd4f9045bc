"""Readers and writers for the reference's binary asset formats (4-byte u32 layout of the shipped files), and the
reference converter's texel-quad baking (slimrt_host_bake_texture).

.mesh (src/slim/serialization/mesh.h:56-122, bvh.h:18-48): header of 5 u32 (vertex, triangle, edge, uv and
normal counts) and the BVH header (node count, height as a 4-byte field whose low byte is the height); then
the AABB, Triangle records (128 B), vertex positions, position indices, edge indices, optional uvs + uv indices,
optional normals + normal indices, and the BVH nodes (32 B).  Trailing bytes after the nodes are ignored
(files written over older, longer ones keep stale tails).

.texture (src/slim/serialization/texture.h:53-70, core/base.h:1174-1282, 1544-1549): the ImageInfo header of 8
u32 (width, height, size, stride, tile width, tile height, mip count, flags), then per mip its width and height
(u32) and (width + 1) * (height + 1) texel quads of 12 bytes.

.bvh (src/slim/serialization/bvh.h:18-48): node count and height (u32), then the nodes (32 B).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .runtime import BVH_NODE, TRIANGLE, lib, _check, _ptr

IMG_LINEAR, IMG_MIPMAP, IMG_WRAP, IMG_NORMAL, IMG_CUBEMAP = 1 << 1, 1 << 4, 1 << 6, 1 << 7, 1 << 8


def read_mesh(path: str) -> dict:
    data = open(path, "rb").read()
    hdr = np.frombuffer(data, "<u4", 7, 0)
    V, T, E, U, N, node_count, height = (int(x) for x in hdr)
    height &= 0xFF
    off = 28

    def take(dtype, count, shape=None):
        nonlocal off
        dt = np.dtype(dtype)
        a = np.frombuffer(data, dt, count, off)
        off += dt.itemsize * count
        return a.reshape(shape) if shape else a

    out = {"vertex_count": V, "triangle_count": T, "edge_count": E, "uvs_count": U, "normals_count": N,
           "node_count": node_count, "height": height}
    out["aabb"] = take("<f4", 6)
    out["triangles"] = take(TRIANGLE, T)
    out["positions"] = take("<f4", 3 * V, (V, 3))
    out["position_indices"] = take("<u4", 3 * T, (T, 3))
    out["edges"] = take("<u4", 2 * E, (E, 2))
    if U:
        out["uvs"] = take("<f4", 2 * U, (U, 2))
        out["uv_indices"] = take("<u4", 3 * T, (T, 3))
    if N:
        out["normals"] = take("<f4", 3 * N, (N, 3))
        out["normal_indices"] = take("<u4", 3 * T, (T, 3))
    out["nodes"] = take(BVH_NODE, node_count)
    out["trailing_bytes"] = len(data) - off
    return out


def write_mesh(path: str, m: dict):
    """The layout read_mesh reads (Mesh save, src/slim/serialization/mesh.h:56-65, 104-122); no trailing bytes."""
    T, V = len(m["triangles"]), len(m["positions"])
    uvs, nrm = m.get("uvs"), m.get("normals")
    U = 0 if uvs is None else len(uvs)
    N = 0 if nrm is None else len(nrm)
    edges = m.get("edges", np.zeros((0, 2), np.uint32))
    parts = [np.array([V, T, len(edges), U, N, len(m["nodes"]), m["height"]], "<u4").tobytes(),
             np.asarray(m["aabb"], "<f4").tobytes(), np.ascontiguousarray(m["triangles"], TRIANGLE).tobytes(),
             np.asarray(m["positions"], "<f4").tobytes(), np.asarray(m["position_indices"], "<u4").tobytes(),
             np.asarray(edges, "<u4").tobytes()]
    if U:
        parts += [np.asarray(uvs, "<f4").tobytes(), np.asarray(m["uv_indices"], "<u4").tobytes()]
    if N:
        parts += [np.asarray(nrm, "<f4").tobytes(), np.asarray(m["normal_indices"], "<u4").tobytes()]
    parts.append(np.ascontiguousarray(m["nodes"], BVH_NODE).tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(parts))


def read_texture(path: str) -> dict:
    data = open(path, "rb").read()
    info = np.frombuffer(data, "<u4", 8, 0)
    out = dict(zip(("width", "height", "size", "stride", "tile_width", "tile_height", "mip_count", "flags"),
                   (int(x) for x in info)))
    off, mips = 32, []
    for _ in range(out["mip_count"]):
        w, h = (int(x) for x in np.frombuffer(data, "<u4", 2, off))
        off += 8
        n = (w + 1) * (h + 1) * 12
        mips.append((w, h, np.frombuffer(data, np.uint8, n, off).reshape(h + 1, w + 1, 12)))
        off += n
    out["mips"] = mips
    out["trailing_bytes"] = len(data) - off
    return out


def write_texture(path: str, t: dict):
    """Texture save (src/slim/core/base.h:1544-1546, serialization/texture.h:61-68); size and stride default to
    width * height and width, tiles to 0, as the converter leaves them."""
    if t["flags"] & (3 << 24):
        raise ValueError("the .texture format holds texel quads only; this texture is in texel storage")
    w, h = t["width"], t["height"]
    info = [w, h, t.get("size", w * h), t.get("stride", w), t.get("tile_width", 0), t.get("tile_height", 0),
            len(t["mips"]), t["flags"]]
    parts = [np.array(info, "<u4").tobytes()]
    for mw, mh, q in t["mips"]:
        parts += [np.array([mw, mh], "<u4").tobytes(), np.ascontiguousarray(q, np.uint8).tobytes()]
    with open(path, "wb") as f:
        f.write(b"".join(parts))


def read_bvh(path: str):
    data = open(path, "rb").read()
    count, height = (int(x) for x in np.frombuffer(data, "<u4", 2, 0))
    return np.frombuffer(data, BVH_NODE, count, 8).copy(), height


def write_bvh(path: str, nodes: np.ndarray, height: int):
    with open(path, "wb") as f:
        f.write(np.array([len(nodes), height], "<u4").tobytes() + np.ascontiguousarray(nodes, BVH_NODE).tobytes())


def bake_texture(rgb: np.ndarray, flags: int) -> dict:
    """The reference converter's texture from an (h, w, 3) u8 image (slimrt_host_bake_texture): a dict in
    read_texture's form."""
    img = np.ascontiguousarray(rgb, np.uint8)
    h, w, _ = img.shape
    total, mips_n, fl = C.c_size_t(), C.c_uint32(), C.c_uint32()
    dims = np.zeros(32, np.uint32)
    fn = lib().slimrt_host_bake_texture
    _check(fn(_ptr(img), w, h, flags, None, 0, C.byref(total), C.byref(mips_n), _ptr(dims), C.byref(fl)), what="bake")
    quads = np.zeros(total.value * 12, np.uint8)
    _check(fn(_ptr(img), w, h, flags, _ptr(quads), total.value, C.byref(total), C.byref(mips_n), _ptr(dims),
              C.byref(fl)), what="bake")
    mips, off = [], 0
    for i in range(mips_n.value):
        mw, mh = int(dims[2 * i]), int(dims[2 * i + 1])
        n = (mw + 1) * (mh + 1) * 12
        mips.append((mw, mh, quads[off:off + n].reshape(mh + 1, mw + 1, 12)))
        off += n
    return {"width": w, "height": h, "size": w * h, "stride": w, "tile_width": 0, "tile_height": 0,
            "mip_count": mips_n.value, "flags": fl.value, "mips": mips}
