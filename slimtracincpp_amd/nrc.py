"""ctypes bindings for the neural radiance cache (include/slimrt_nrc.h).  Buffers are torch tensors on the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import runtime

N_PARAMS = 19728
N_RAW = 14

_bound = False


def _lib():
    global _bound
    L = runtime.lib()
    if not _bound:
        vp, u32, f32, P = C.c_void_p, C.c_uint32, C.c_float, C.POINTER
        sig = {"slimrt_nrc_create": [P(vp), u32, vp, vp], "slimrt_nrc_destroy": [vp], "slimrt_nrc_set_stream": [vp, vp],
               "slimrt_nrc_get_params": [vp, vp], "slimrt_nrc_set_params": [vp, vp],
               "slimrt_nrc_infer": [vp, vp, u32, vp], "slimrt_nrc_gradients": [vp, vp, vp, u32, vp, vp],
               "slimrt_nrc_train_step": [vp, vp, vp, u32, f32, vp]}
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
