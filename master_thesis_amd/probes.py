"""Python front end of the irradiance probes (rt_ddgi_* in include/rt.h): the
Vulkan renderer's Majercik2019 probe field as a cache of the path tracer."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .tracer import Tracer


class ProbeField:
    """A probe grid on the context of `tracer`.  count, start, spacing: the
    scene file's Probes entry (probeCount, probeStartCorner,
    betweenProbeDistance); the other arguments are Majercik2019's renderer keys
    (perProbeRayCount, probeSampleSideLength, maxProbeRayDistance,
    depthSharpness, normalBias, energyPreservation, shadowCountProbe,
    linearBlending) plus hysteresis, multi_bounce, rotate and
    reference_visibility (rt_ddgi_config, RT_DDGI_*)."""

    def __init__(self, tracer: Tracer, count, start, spacing, rays: int = 256, side: int = 8,
                 max_distance: float = 1000.0, sharpness: float = 0.8, normal_bias: float = 0.001,
                 energy: float = 1.0, hysteresis: float = 0.0, shadow_count: int = 1, linear: bool = False,
                 multi_bounce: bool = False, rotate: bool = False, reference_visibility: bool = False):
        self.tracer = tracer
        self.lib = tracer.lib
        cfg = L.rt_ddgi_config()
        for k in range(3):
            cfg.probe_count[k] = int(count[k])
            cfg.probe_start[k] = float(start[k])
            cfg.probe_spacing[k] = float(spacing[k])
        cfg.rays_per_probe = rays
        cfg.side = side
        cfg.max_ray_distance = max_distance
        cfg.depth_sharpness = sharpness
        cfg.normal_bias = normal_bias
        cfg.energy_preservation = energy
        cfg.hysteresis = hysteresis
        cfg.shadow_count = shadow_count
        cfg.flags = ((L.RT_DDGI_LINEAR_BLENDING if linear else 0) | (L.RT_DDGI_MULTI_BOUNCE if multi_bounce else 0) |
                     (L.RT_DDGI_ROTATE if rotate else 0) |
                     (L.RT_DDGI_REFERENCE_VISIBILITY if reference_visibility else 0))
        self.cfg = cfg
        h = C.c_void_p()
        tracer._check(self.lib.rt_ddgi_create(tracer.h, C.byref(cfg), C.byref(h)))
        self.h = h
        tracer.caches.add(self)
        w, hh = C.c_uint32(), C.c_uint32()
        self._check(self.lib.rt_ddgi_atlas_size(self.h, C.byref(w), C.byref(hh)))
        self.atlas_shape = (hh.value, w.value)
        self.num_probes = int(count[0]) * int(count[1]) * int(count[2])
        self.rays = rays

    def close(self):
        if getattr(self, "h", None):
            self.lib.rt_ddgi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        self.tracer._check(rc)

    def update(self, settings) -> float:
        """One field update with the tracer's scene; returns its device ms."""
        self._check(self.lib.rt_ddgi_update(self.h, C.byref(settings)))
        return self.tracer.timing()["last_call_ms"]

    def update_surfels(self, surfels: np.ndarray) -> None:
        s = np.ascontiguousarray(surfels, np.float32)
        assert s.size == self.num_probes * self.rays * L.RT_DDGI_SURFEL_FLOATS
        self._check(self.lib.rt_ddgi_update_surfels(self.h, s.ctypes.data, L.RT_PTR_HOST))

    def surfels(self):
        """(surfels (probes * rays, 8), rotation (3, 3)) of the last update."""
        s = np.zeros((self.num_probes * self.rays, L.RT_DDGI_SURFEL_FLOATS), np.float32)
        r = np.zeros(9, np.float32)
        self._check(self.lib.rt_ddgi_surfels(self.h, s.ctypes.data, r.ctypes.data, L.RT_PTR_HOST))
        return s, r.reshape(3, 3)

    def atlas(self):
        """(irradiance (H, W, 4), depth (H, W, 2))."""
        H, W = self.atlas_shape
        irr = np.zeros((H, W, 4), np.float32)
        dep = np.zeros((H, W, 2), np.float32)
        self._check(self.lib.rt_ddgi_get_atlas(self.h, irr.ctypes.data, dep.ctypes.data, L.RT_PTR_HOST))
        return irr, dep

    def set_atlas(self, irr: np.ndarray, dep: np.ndarray) -> None:
        irr = np.ascontiguousarray(irr, np.float32)
        dep = np.ascontiguousarray(dep, np.float32)
        H, W = self.atlas_shape
        assert irr.size == H * W * 4 and dep.size == H * W * 2
        self._check(self.lib.rt_ddgi_set_atlas(self.h, irr.ctypes.data, dep.ctypes.data, L.RT_PTR_HOST))

    def eval(self, samples: np.ndarray) -> np.ndarray:
        """Lookups for samples (n, 9): position, normal, incoming direction."""
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, 9)
        out = np.zeros((s.shape[0], 3), np.float32)
        self._check(self.lib.rt_ddgi_eval(self.h, s.ctypes.data, s.shape[0], out.ctypes.data, L.RT_PTR_HOST))
        return out

    def render(self, cams, settings, query_vertex: int = 0, out=None):
        """rt_render_ddgi for a camera or a list of cameras.  Returns (rgb,
        stats) like RadianceCache.render."""
        single = not isinstance(cams, (list, tuple))
        cl = [cams] if single else list(cams)
        V = len(cl)
        carr = (L.rt_camera * V)(*cl)
        H, W = cl[0].height, cl[0].width
        st = L.rt_nrc_stats()
        dev = out is not None and L._is_device(out)
        if out is None:
            out = np.zeros((V, H, W, 3), np.float32)
        self._check(self.lib.rt_render_ddgi(self.tracer.h, self.h, carr, V, C.byref(settings), query_vertex,
                                            L._ptr(out), None, L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST,
                                            C.byref(st)))
        if single and not dev:
            out = out[0]
        return out, st
