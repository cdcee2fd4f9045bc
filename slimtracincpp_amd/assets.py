"""Readers for the reference's binary asset formats (4-byte u32 layout of the shipped files).

.mesh (src/slim/serialization/mesh.h:56-122, bvh.h:18-48): header of 5 u32 (vertex, triangle, edge, uv and
normal counts) and the BVH header (node count, height as a 4-byte field whose low byte is the height); then
the AABB, Triangle records (128 B), vertex positions, position indices, edge indices, optional uvs + uv indices,
optional normals + normal indices, and the BVH nodes (32 B).  Trailing bytes after the nodes are ignored
(files written over older, longer ones keep stale tails).
"""
from __future__ import annotations

import numpy as np

from .runtime import BVH_NODE, TRIANGLE


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
