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
        tracer.caches.add(self)

    def close(self):
        if getattr(self, "h", None):
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

    def last_records(self):
        """(feats (n, 9), targets (n, 3)) of the last render's training records."""
        n = C.c_uint32(0)
        self._check(self.lib.rt_render_nrc_records(self.tracer.h, None, None, C.byref(n)))
        f = np.zeros((n.value, 9), np.float32)
        t = np.zeros((n.value, 3), np.float32)
        self._check(self.lib.rt_render_nrc_records(self.tracer.h, L._ptr(f), L._ptr(t), C.byref(n)))
        return f, t

    def render(self, cams, settings, query_vertex: int = 1, train_every: int = 16, batch: int = 16384,
               max_steps: int = 4, train: bool = True, query: bool = True, nee: bool = True, out=None):
        """One frame of rt_render_nrc for a camera or a list of cameras.
        train_every = 0: no training samples.  Returns (rgb, stats): rgb
        (H, W, 3) or (V, H, W, 3) float32 (numpy, or the torch tensor `out`
        on the device), stats an rt_nrc_stats."""
        single = not isinstance(cams, (list, tuple))
        cl = [cams] if single else list(cams)
        V = len(cl)
        carr = (L.rt_camera * V)(*cl)
        H, W = cl[0].height, cl[0].width
        o = L.rt_nrc_render(query_vertex, train_every, batch, max_steps,
                            (0 if train else L.RT_NRC_NO_TRAIN) | (0 if query else L.RT_NRC_NO_QUERY) |
                            (L.RT_NRC_NO_RECORDS if train_every == 0 else 0) | (0 if nee else L.RT_NRC_NO_NEE))
        st = L.rt_nrc_stats()
        dev = out is not None and L._is_device(out)
        if out is None:
            out = np.zeros((V, H, W, 3), np.float32)
        self._check(self.lib.rt_render_nrc(self.tracer.h, self.h, carr, V, C.byref(settings), C.byref(o),
                                           L._ptr(out), None, L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST,
                                           C.byref(st)))
        if single and not dev:
            out = out[0]
        return out, st
