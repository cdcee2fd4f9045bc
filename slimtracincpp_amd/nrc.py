"""ctypes bindings for the neural radiance cache (include/slimrt_nrc.h).  Buffers are torch tensors on the GPU.

NeuralCache wraps the network (inference, gradients, Adam steps); CachedFrames renders beauty frames through a
Renderer with the cache answering after `cache_depth` bounces, training it online from full-path training waves.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import runtime

N_PARAMS = 19728
N_RAW = 14

_bound = False


class NrcRenderParams(C.Structure):
    _fields_ = [("cache_depth", C.c_uint32), ("train_one_in", C.c_uint32), ("train_steps", C.c_uint32),
                ("train_batch", C.c_uint32), ("learning_rate", C.c_float), ("gate_threshold", C.c_float),
                ("gate_min_frames", C.c_uint32), ("gate_cell_size", C.c_float), ("gate_cells_log2", C.c_uint32),
                ("frame", C.c_uint32), ("flags", C.c_uint32), ("reserved", C.c_uint32)]


RESET_GATE = 1


def _lib():
    global _bound
    L = runtime.lib()
    if not _bound:
        vp, u32, f32, P = C.c_void_p, C.c_uint32, C.c_float, C.POINTER
        sig = {"slimrt_nrc_create": [P(vp), u32, vp, vp], "slimrt_nrc_destroy": [vp], "slimrt_nrc_set_stream": [vp, vp],
               "slimrt_nrc_get_params": [vp, vp], "slimrt_nrc_set_params": [vp, vp],
               "slimrt_nrc_get_state": [vp, vp, vp, vp, P(u32)], "slimrt_nrc_set_state": [vp, vp, vp, vp, u32],
               "slimrt_nrc_infer": [vp, vp, u32, vp], "slimrt_nrc_gradients": [vp, vp, vp, u32, vp, vp],
               "slimrt_nrc_train_step": [vp, vp, vp, u32, f32, vp],
               "slimrt_nrc_infer_queue": [vp, vp, vp, u32, vp],
               "slimrt_nrc_train_queue": [vp, vp, vp, vp, u32, u32, u32, f32, vp],
               "slimrt_nrc_frame_stats": [vp, vp],
               "slimrt_nrc_frame_samples": [vp, vp, vp, vp, vp, u32, P(u32)],
               "slimrt_render_nrc": [vp, P(runtime.RenderParams), vp, P(NrcRenderParams), vp, vp]}
        for name, args in sig.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_int
        L.slimrt_nrc_last_error.argtypes = [vp]
        L.slimrt_nrc_last_error.restype = C.c_char_p
        _bound = True
    return L


class NeuralCache:
    def __init__(self, center, inv_half, seed: int = 0):
        self.h = C.c_void_p()
        c = np.ascontiguousarray(center, np.float32)
        ih = np.ascontiguousarray(inv_half, np.float32)
        rc = _lib().slimrt_nrc_create(C.byref(self.h), seed, c.ctypes.data_as(C.c_void_p), ih.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise runtime.SlimrtError(rc, "slimrt_nrc_create failed")
        self.center, self.inv_half = c, ih

    def _check(self, rc):
        if rc != 0:
            raise runtime.SlimrtError(rc, _lib().slimrt_nrc_last_error(self.h).decode())

    def close(self):
        if self.h:
            _lib().slimrt_nrc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        self._check(_lib().slimrt_nrc_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None))

    def params(self) -> np.ndarray:
        p = np.zeros(N_PARAMS, np.float32)
        self._check(_lib().slimrt_nrc_get_params(self.h, p.ctypes.data_as(C.c_void_p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == N_PARAMS
        self._check(_lib().slimrt_nrc_set_params(self.h, p.ctypes.data_as(C.c_void_p)))

    def state(self) -> dict:
        """The training state for a checkpoint: params, Adam moments m and v, step count."""
        p, m, v = (np.zeros(N_PARAMS, np.float32) for _ in range(3))
        step = C.c_uint32()
        self._check(_lib().slimrt_nrc_get_state(self.h, p.ctypes.data_as(C.c_void_p), m.ctypes.data_as(C.c_void_p),
                                                v.ctypes.data_as(C.c_void_p), C.byref(step)))
        return {"params": p, "m": m, "v": v, "step": step.value}

    def set_state(self, st: dict):
        p, m, v = (np.ascontiguousarray(st[k], np.float32) for k in ("params", "m", "v"))
        assert p.size == m.size == v.size == N_PARAMS
        self._check(_lib().slimrt_nrc_set_state(self.h, p.ctypes.data_as(C.c_void_p), m.ctypes.data_as(C.c_void_p),
                                                v.ctypes.data_as(C.c_void_p), int(st["step"])))

    def infer(self, raw, out):
        """raw: (n, 14) float32 cuda tensor; out: (n, 3) float32 cuda tensor."""
        n = raw.shape[0]
        self._check(_lib().slimrt_nrc_infer(self.h, C.c_void_p(raw.data_ptr()), n, C.c_void_p(out.data_ptr())))

    def gradients(self, raw, targets):
        g = np.zeros(N_PARAMS, np.float32)
        loss = C.c_float()
        self._check(_lib().slimrt_nrc_gradients(self.h, C.c_void_p(raw.data_ptr()), C.c_void_p(targets.data_ptr()),
                                                raw.shape[0], g.ctypes.data_as(C.c_void_p), C.byref(loss)))
        return g, loss.value

    def train_step(self, raw, targets, lr: float = 5e-3, want_loss: bool = False):
        loss = C.c_float()
        self._check(_lib().slimrt_nrc_train_step(self.h, C.c_void_p(raw.data_ptr()), C.c_void_p(targets.data_ptr()),
                                                 raw.shape[0], float(lr), C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def infer_queue(self, raw, count, out):
        """Inference on the first min(count[0], raw.shape[0]) records; count is a uint32 (int32) cuda tensor."""
        self._check(_lib().slimrt_nrc_infer_queue(self.h, C.c_void_p(raw.data_ptr()), C.c_void_p(count.data_ptr()),
                                                  raw.shape[0], C.c_void_p(out.data_ptr())))

    def train_queue(self, raw, targets, count, steps: int, batch: int, lr: float = 5e-3, losses=None):
        self._check(_lib().slimrt_nrc_train_queue(self.h, C.c_void_p(raw.data_ptr()), C.c_void_p(targets.data_ptr()),
                                                  C.c_void_p(count.data_ptr()), raw.shape[0], steps, batch, float(lr),
                                                  C.c_void_p(losses.data_ptr()) if losses is not None else None))


def scene_box(scene: runtime.SceneBuild):
    """Centre and inverse half extent of the scene BVH's root box (the cache's position normalisation)."""
    root = scene.scene_nodes[0]
    lo = np.asarray(root["min"], np.float64)
    hi = np.asarray(root["max"], np.float64)
    center = (lo + hi) * 0.5
    half = np.maximum((hi - lo) * 0.5, 1e-3)
    return center.astype(np.float32), (1.0 / half).astype(np.float32)


class CachedFrames:
    """Beauty frames rendered with the neural radiance cache (slimrt_render_nrc), frame after frame."""

    def __init__(self, renderer: runtime.Renderer, cache: NeuralCache | None = None, seed: int = 0, cache_depth: int = 2,
                 train_one_in: int = 8, train_steps: int = 4, train_batch: int = 16384, learning_rate: float = 5e-3,
                 gate_threshold: float = 0.25, gate_min_frames: int = 4, gate_cell_size: float = 0.25,
                 gate_cells_log2: int = 20):
        self.r = renderer
        if cache is None:
            center, inv_half = scene_box(renderer.scene)
            cache = NeuralCache(center, inv_half, seed)
        self.cache = cache
        self.settings = NrcRenderParams(cache_depth, train_one_in, train_steps, train_batch, learning_rate,
                                        gate_threshold, gate_min_frames, gate_cell_size, gate_cells_log2, 0, RESET_GATE, 0)

    def render_device(self, params: runtime.RenderParams, pixels_ptr: int, depths_ptr: int, **overrides):
        st = NrcRenderParams.from_buffer_copy(self.settings)
        for k, v in overrides.items():
            setattr(st, k, v)
        rc = _lib().slimrt_render_nrc(self.r.ctx, C.byref(params), self.cache.h, C.byref(st), C.c_void_p(pixels_ptr),
                                      C.c_void_p(depths_ptr))
        runtime._check(rc, self.r.ctx)
        self.settings.frame += 1
        self.settings.flags = 0

    def stats(self) -> dict:
        s = np.zeros(8, np.uint32)
        runtime._check(_lib().slimrt_nrc_frame_stats(self.r.ctx, s.ctypes.data_as(C.c_void_p)), self.r.ctx)
        f = s.view(np.float32)
        return {"queries": int(s[0]), "samples": int(s[1]), "fallbacks": int(s[2]), "train_pixels": int(s[3]),
                "loss_first": float(f[4]), "loss_last": float(f[5])}

    def samples(self, cap: int | None = None) -> dict:
        """The last frame's training samples: raw (n, 14), target (n, 3), pixel (canvas offset), vertex."""
        L = _lib()
        n = C.c_uint32()
        runtime._check(L.slimrt_nrc_frame_samples(self.r.ctx, None, None, None, None, 0, C.byref(n)), self.r.ctx)
        m = n.value if cap is None else min(cap, n.value)
        raw = np.zeros((m, N_RAW), np.float32)
        tgt = np.zeros((m, 3), np.float32)
        pix = np.zeros(m, np.uint32)
        vtx = np.zeros(m, np.uint32)
        if m:
            vp = C.c_void_p
            runtime._check(L.slimrt_nrc_frame_samples(self.r.ctx, vp(raw.ctypes.data), vp(tgt.ctypes.data),
                                                      vp(pix.ctypes.data), vp(vtx.ctypes.data), m, C.byref(n)),
                           self.r.ctx)
        return {"raw": raw, "target": tgt, "pixel": pix, "vertex": vtx, "count": n.value}
