"""Python front end of the C ABI: a context that owns one scene, its BVH and
the render / trace entry points.  Host arrays are numpy; device arrays are
torch CUDA (HIP) tensors, passed by pointer (no copies)."""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib as L
from .scenes import Scene


class Tracer:
    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.rt_create(C.byref(h), device)
        if rc != L.RT_OK:
            raise L.RTError(rc, self.lib.rt_last_error(None).decode())
        self.h = h
        self.scene: Scene | None = None
        self.info = None
        self.caches = weakref.WeakSet()  # radiance caches and probe fields on this context, closed first

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        for c in list(self.caches):
            c.close()
        if self.h:
            self.lib.rt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != L.RT_OK:
            raise L.RTError(rc, self.lib.rt_last_error(self.h).decode())

    # -- scene --------------------------------------------------------------
    def upload(self, scene: Scene) -> None:
        scene.validate()
        self._keep = (np.ascontiguousarray(scene.pos), np.ascontiguousarray(scene.nrm),
                      np.ascontiguousarray(scene.idx), np.ascontiguousarray(scene.tri_mat),
                      np.ascontiguousarray(scene.materials))
        pos, nrm, idx, tm, mat = self._keep
        d = L.rt_scene_desc(pos.ctypes.data, nrm.ctypes.data, idx.ctypes.data, tm.ctypes.data, mat.ctypes.data,
                            pos.shape[0], idx.shape[0], mat.shape[0])
        self._check(self.lib.rt_scene_upload(self.h, C.byref(d)))
        self.scene = scene

    def set_stream(self, stream_ptr: int | None) -> None:
        self._check(self.lib.rt_set_stream(self.h, stream_ptr))

    def set_material_colors(self, rgb) -> None:
        kind = L.RT_PTR_DEVICE if L._is_device(rgb) else L.RT_PTR_HOST
        if kind == L.RT_PTR_HOST:
            rgb = np.ascontiguousarray(rgb, np.float32)
        self._check(self.lib.rt_set_material_colors(self.h, L._ptr(rgb), kind))

    def build(self, leaf_size: int = 4, cluster: bool = False) -> dict:
        """Builds the BVH4: the hierarchy of the Morton order, or with
        cluster=True the clustered one (RT_BUILD_CLUSTER)."""
        opts = L.rt_build_opts(leaf_size, L.RT_BUILD_CLUSTER if cluster else 0)
        info = L.rt_bvh_info()
        self._check(self.lib.rt_build(self.h, C.byref(opts), C.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in L.rt_bvh_info._fields_}
        return self.info

    def export_bvh(self):
        nn, nt = C.c_uint32(0), C.c_uint32(0)
        self._check(self.lib.rt_bvh_export(self.h, None, C.byref(nn), None, C.byref(nt)))
        nodes = np.zeros(nn.value, L.NODE_DTYPE)
        tris = np.zeros(nt.value, L.TRI_DTYPE)
        self._check(self.lib.rt_bvh_export(self.h, nodes.ctypes.data, C.byref(nn), tris.ctypes.data, C.byref(nt)))
        return nodes, tris

    def export_traversal(self) -> np.ndarray:
        """The traversal's 64-byte quantised nodes (QNODE_DTYPE), node order
        as export_bvh."""
        nn = C.c_uint32(0)
        self._check(self.lib.rt_bvh_export_traversal(self.h, None, C.byref(nn)))
        q = np.zeros(nn.value, L.QNODE_DTYPE)
        self._check(self.lib.rt_bvh_export_traversal(self.h, q.ctypes.data, C.byref(nn)))
        return q

    # -- queries ------------------------------------------------------------
    def trace(self, rays, hits=None, counters: bool = False):
        """Closest hit.  rays: numpy RAY_DTYPE array (host) or a contiguous
        CUDA tensor of n*8 float32 (device).  Returns hits (and counters)."""
        dev = L._is_device(rays)
        if dev:
            n = rays.numel() // 8
            if hits is None:
                import torch
                hits = torch.empty(n * 4, dtype=torch.int32, device=rays.device)
        else:
            rays = np.ascontiguousarray(rays)
            assert rays.dtype == L.RAY_DTYPE
            n = rays.shape[0]
            if hits is None:
                hits = np.zeros(n, L.HIT_DTYPE)
        cnt = L.rt_trace_counters()
        self._check(self.lib.rt_trace_closest(self.h, L._ptr(rays), n, L._ptr(hits),
                                              L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST,
                                              C.byref(cnt) if counters else None))
        if counters:
            return hits, {f: getattr(cnt, f) for f, _ in L.rt_trace_counters._fields_}
        return hits

    def trace_noalloc(self, rays, hits, status=None) -> None:
        """Closest hit through rt_trace_closest_noalloc: rays (n*8 float32)
        and hits (n*4 int32) are CUDA tensors the caller owns, status an
        optional one-word int32 CUDA tensor; enqueued on the context stream,
        nothing is allocated."""
        assert L._is_device(rays) and L._is_device(hits), "device tensors only"
        n = rays.numel() // 8
        assert hits.numel() >= 4 * n
        self._check(self.lib.rt_trace_closest_noalloc(self.h, L._ptr(rays), n, L._ptr(hits),
                                                      None if status is None else L._ptr(status)))

    def trace_any(self, rays, out=None):
        """Occlusion flags (uint8, 1 = some triangle within (tmin, tmax)).
        rays: numpy RAY_DTYPE array or a CUDA tensor of n*8 float32."""
        dev = L._is_device(rays)
        if dev:
            n = rays.numel() // 8
            if out is None:
                import torch
                out = torch.empty(n, dtype=torch.uint8, device=rays.device)
        else:
            rays = np.ascontiguousarray(rays)
            assert rays.dtype == L.RAY_DTYPE
            n = rays.shape[0]
            if out is None:
                out = np.zeros(n, np.uint8)
        self._check(self.lib.rt_trace_any(self.h, L._ptr(rays), n, L._ptr(out),
                                          L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST))
        return out

    def occluded(self, start, end):
        """Scene::isOccluded for segment arrays start, end of shape (n, 3)."""
        seg = np.ascontiguousarray(np.concatenate([np.asarray(start, np.float32).reshape(-1, 3),
                                                   np.asarray(end, np.float32).reshape(-1, 3)], axis=1))
        out = np.zeros(seg.shape[0], np.uint8)
        self._check(self.lib.rt_occluded(self.h, L._ptr(seg), seg.shape[0], L._ptr(out), L.RT_PTR_HOST))
        return out

    def timing(self) -> dict:
        t = L.rt_timing()
        self._check(self.lib.rt_get_timing(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in L.rt_timing._fields_}

    # -- rendering ----------------------------------------------------------
    def render(self, cam: L.rt_camera, settings: L.rt_settings, rgb=None, rgb8=None, want_rgb8: bool = False):
        """Returns (rgb float32 HxWx3, rgb8 uint8 HxWx3 or None).  Pass torch
        CUDA tensors as rgb/rgb8 to render straight into device memory."""
        W, H = cam.width, cam.height
        dev = rgb is not None and L._is_device(rgb)
        if not dev:
            rgb = np.zeros((H, W, 3), np.float32) if rgb is None else rgb
            if want_rgb8 and rgb8 is None:
                rgb8 = np.zeros((H, W, 3), np.uint8)
        self._check(self.lib.rt_render(self.h, C.byref(cam), C.byref(settings), L._ptr(rgb),
                                       L._ptr(rgb8) if rgb8 is not None else None,
                                       L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST))
        return rgb, rgb8

    # -- post-filters ---------------------------------------------------
    def _filter(self, fn, img, *args):
        dev = L._is_device(img)
        if dev:
            import torch
            img = img.contiguous()
            out = torch.empty_like(img)
        else:
            img = np.ascontiguousarray(img, np.float32)
            out = np.empty_like(img)
        H, W = img.shape[0], img.shape[1]
        self._check(fn(self.h, L._ptr(img), L._ptr(out), W, H, *args, L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST))
        return out

    def denoise_gauss(self, img, kernel_size: int, sigma: float):
        """GaussDenoiser on an (H, W, 3) float image (numpy or device tensor)."""
        return self._filter(self.lib.rt_denoise_gauss, img, kernel_size, sigma)

    def denoise_median(self, img):
        """MedianDenoiser (3x3 median per channel) on an (H, W, 3) float image."""
        return self._filter(self.lib.rt_denoise_median, img)

    def render_bdpt(self, cam: L.rt_camera, spp: int, vision_jumps: int = 5, light_jumps: int = 5,
                    max_depth: int = 5, seed: int = 1, sample_offset: int = 0, want_rgb8: bool = False):
        """BidirectionalPathTracer frame; returns (rgb HxWx3, rgb8 or None)."""
        s = L.rt_bdpt_settings(vision_jumps, light_jumps, max_depth, spp, seed, sample_offset, 0)
        rgb = np.zeros((cam.height, cam.width, 3), np.float32)
        rgb8 = np.zeros((cam.height, cam.width, 3), np.uint8) if want_rgb8 else None
        self._check(self.lib.rt_render_bdpt(self.h, C.byref(cam), C.byref(s), L._ptr(rgb),
                                            L._ptr(rgb8) if rgb8 is not None else None, L.RT_PTR_HOST))
        return rgb, rgb8

    def render_views(self, cams, settings: L.rt_settings, rgb=None, want_rgb8: bool = False):
        """Renders len(cams) views in one launch sequence; returns
        (rgb (V, H, W, 3), rgb8 or None).  rgb may be a CUDA tensor."""
        V = len(cams)
        arr = (L.rt_camera * V)(*cams)
        W, H = cams[0].width, cams[0].height
        dev = rgb is not None and L._is_device(rgb)
        rgb8 = None
        if not dev:
            rgb = np.zeros((V, H, W, 3), np.float32) if rgb is None else rgb
            if want_rgb8:
                rgb8 = np.zeros((V, H, W, 3), np.uint8)
        self._check(self.lib.rt_render_views(self.h, arr, V, C.byref(settings), L._ptr(rgb),
                                             L._ptr(rgb8) if rgb8 is not None else None,
                                             L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST))
        return rgb, rgb8

    def render_adjoint(self, cam: L.rt_camera, settings: L.rt_settings, d_image, d_color=None):
        dev = L._is_device(d_image)
        M = self.scene.materials.shape[0]
        if d_color is None:
            if dev:
                import torch
                d_color = torch.zeros((M, 3), dtype=torch.float32, device=d_image.device)
            else:
                d_color = np.zeros((M, 3), np.float32)
        if not dev:
            d_image = np.ascontiguousarray(d_image, np.float32)
        self._check(self.lib.rt_render_adjoint(self.h, C.byref(cam), C.byref(settings), L._ptr(d_image),
                                               L._ptr(d_color), L.RT_PTR_DEVICE if dev else L.RT_PTR_HOST))
        return d_color


def make_settings(spp: int, max_depth: int, seed: int = 1, flags: int = 0, sample_offset: int = 0,
                  sky=(0.0, 0.0, 0.0), sun_dir=(0.0, 1.0, 0.0), sun_radius_deg: float = 0.0,
                  sun_radiance=(0.0, 0.0, 0.0)) -> L.rt_settings:
    s = L.rt_settings()
    s.max_depth = max_depth
    s.spp = spp
    s.seed = seed
    s.flags = flags
    s.sample_offset = sample_offset
    sd = np.asarray(sun_dir, np.float64)
    sd = sd / np.linalg.norm(sd)
    for k in range(3):
        s.sky[k] = sky[k]
        s.sun_dir[k] = sd[k]
        s.sun_radiance[k] = sun_radiance[k]
    s.sun_cos = float(np.cos(np.radians(sun_radius_deg)))
    s.num_views = 1
    return s
