"""Python front end of the learned radiance cache (rt_nrc_* in include/rt.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .tracer import Tracer


class RadianceCache:
    """MLP radiance cache on the context of `tracer` (after its scene upload,
    whose bounds normalise positions unless `bounds` is given)."""

    def __init__(self, tracer: Tracer, lr: float = 1e-2, seed: int = 0, bounds=None):
        self.tracer = tracer
        self.lib = tracer.lib
        cfg = L.rt_nrc_config()
        cfg.lr = lr
        cfg.seed = seed
        if bounds is not None:
            for k in range(3):
                cfg.bounds_lo[k] = bounds[0][k]
                cfg.bounds_hi[k] = bounds[1][k]
        h = C.c_void_p()
        tracer._check(self.lib.rt_nrc_create(tracer.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.rt_nrc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        self.tracer._check(rc)

    def num_params(self) -> int:
        n = C.c_uint32(0)
        self._check(self.lib.rt_nrc_params(self.h, None, C.byref(n), 0))
        return n.value

    def params(self) -> np.ndarray:
        n = C.c_uint32(self.num_params())
        p = np.zeros(n.value, np.float32)
        self._check(self.lib.rt_nrc_params(self.h, p.ctypes.data, C.byref(n), 0))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        n = C.c_uint32(p.shape[0])
        self._check(self.lib.rt_nrc_params(self.h, p.ctypes.data, C.byref(n), 1))

    def infer(self, feats, out=None):
        dev = L._is_device(feats)
        if not dev:
            feats = np.ascontiguousarray(feats, np.float32)
        n = (feats.numel() if dev else feats.size) // 9
        if out is None:
            if dev:
                import torch
                out = torch.empty((n, 3), dtype=torch.float32, device=feats.device)
            else:
                out = np.zeros((n, 3), np.float32)
        self._check(self.lib.rt_nrc_infer(self.h, L._ptr(feats), n, L._ptr(out),
                                          L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST))
        return out

    def gradient(self, feats, targets):
        feats = np.ascontiguousarray(feats, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        g = np.zeros(self.num_params(), np.float32)
        loss = C.c_float()
        self._check(self.lib.rt_nrc_gradient(self.h, feats.ctypes.data, targets.ctypes.data, feats.shape[0],
                                             g.ctypes.data, C.byref(loss), L.RT_PTR_HOST))
        return g, loss.value

    def train(self, feats, targets, want_loss: bool = True):
        dev = L._is_device(feats)
        if not dev:
            feats = np.ascontiguousarray(feats, np.float32)
            targets = np.ascontiguousarray(targets, np.float32)
        n = (feats.numel() if dev else feats.size) // 9
        loss = C.c_float()
        self._check(self.lib.rt_nrc_train(self.h, L._ptr(feats), L._ptr(targets), n,
                                          C.byref(loss) if want_loss else None,
                                          L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST))
        return loss.value if want_loss else None
