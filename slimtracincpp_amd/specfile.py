"""Binary container for a `SceneSpec`, read by the CPU oracle and by the compiled reference driver.

Layout (little-endian):
    magic  8 bytes  b"SLMSPEC1"
    u32    section count
    u32    reserved (0)
    per section:
        char[24] name (NUL padded)
        u32      dtype  (0 f32, 1 u32, 2 i32, 3 u8)
        u32      element count
        bytes    payload, zero-padded to a multiple of 8

Section names and meanings are those written by `write_spec` below; the C reader is oracle/specfile.h.
"""
from __future__ import annotations

import struct

import numpy as np

from .scenes import SceneSpec

_DTYPES = {np.dtype(np.float32): 0, np.dtype(np.uint32): 1, np.dtype(np.int32): 2, np.dtype(np.uint8): 3}


def _sections(spec: SceneSpec):
    n_light = 0 if spec.light_f is None else int(spec.light_f.shape[0])
    frame = np.array([spec.width, spec.height, int(spec.ssaa), spec.max_depth, spec.render_mode,
                      spec.skybox_color, spec.skybox_radiance, spec.skybox_irradiance,
                      spec.geometry_count, int(spec.mat_f.shape[0]), n_light,
                      len(spec.meshes), len(spec.textures)], np.int32)
    yield "frame", frame
    cam = np.concatenate([spec.camera_euler, spec.camera_position,
                          np.array([spec.focal_length], np.float32)]).astype(np.float32)
    yield "camera", cam
    yield "geo_f", spec.geo_f
    yield "geo_u", spec.geo_u
    yield "mat_f", spec.mat_f
    yield "mat_u", spec.mat_u
    if n_light:
        yield "light_f", spec.light_f
        yield "light_u", spec.light_u
    for k, m in enumerate(spec.meshes):
        n_nrm = 0 if m.normals is None else int(m.normals.shape[0])
        n_uv = 0 if m.uvs is None else int(m.uvs.shape[0])
        yield f"mesh{k}_info", np.array([m.positions.shape[0], m.triangle_count, n_nrm, n_uv, m.leaf_size],
                                        np.uint32)
        yield f"mesh{k}_pos", m.positions
        yield f"mesh{k}_pidx", m.position_indices
        if n_nrm:
            yield f"mesh{k}_nrm", m.normals
            yield f"mesh{k}_nidx", m.normal_indices
        if n_uv:
            yield f"mesh{k}_uv", m.uvs
            yield f"mesh{k}_uidx", m.uv_indices
        if m.edges is not None:
            yield f"mesh{k}_edges", np.ascontiguousarray(m.edges, np.uint32).reshape(-1)
    for k, t in enumerate(spec.textures):
        yield f"tex{k}_info", np.array([t.width, t.height, t.flags, len(t.mips)], np.uint32)
        yield f"tex{k}_dims", np.array(t.mip_dims, np.uint32).reshape(-1)
        for mi, q in enumerate(t.mips):
            q = np.ascontiguousarray(q)
            yield f"tex{k}_mip{mi}", q.view(np.uint8) if q.dtype == np.float16 else q


def spec_bytes(spec: SceneSpec) -> bytes:
    parts = []
    count = 0
    for name, arr in _sections(spec):
        arr = np.ascontiguousarray(arr)
        dt = _DTYPES[arr.dtype]
        payload = arr.tobytes()
        pad = (-len(payload)) % 8
        nm = name.encode()
        assert len(nm) < 24, name
        parts.append(struct.pack("<24sII", nm, dt, arr.size) + payload + b"\0" * pad)
        count += 1
    return b"SLMSPEC1" + struct.pack("<II", count, 0) + b"".join(parts)


def write_spec(spec: SceneSpec, path: str) -> str:
    with open(path, "wb") as f:
        f.write(spec_bytes(spec))
    return path


_NP = {0: np.float32, 1: np.uint32, 2: np.int32, 3: np.uint8}


def read_container(path_or_bytes) -> dict:
    """Read any file in this container format (specs, and the result files of the oracle tools)."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as f:
            data = f.read()
    if data[:8] != b"SLMSPEC1":
        raise ValueError("not a SLMSPEC1 container")
    (count, _) = struct.unpack_from("<II", data, 8)
    off = 16
    out = {}
    for _ in range(count):
        name, dt, n = struct.unpack_from("<24sII", data, off)
        off += 32
        dtype = np.dtype(_NP[dt])
        nbytes = n * dtype.itemsize
        out[name.rstrip(b"\0").decode()] = np.frombuffer(data, dtype, n, off).copy()
        off += nbytes + ((-nbytes) % 8)
    return out
