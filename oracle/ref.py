"""ctypes binding of oracle/_ref/libref.so: the reference software renderer
compiled from its own sources (oracle/Makefile target `ref`).

TEST INFRASTRUCTURE, available only where the reference sources were present
at build time; callers skip when `available()` is False.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import REF_PATH

_lib = None


def available() -> bool:
    return REF_PATH.exists()


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        lib = C.CDLL(str(REF_PATH))
        P = C.c_void_p
        lib.ref_intersect.argtypes = [P] * 7
        lib.ref_intersect.restype = C.c_int
        lib.ref_scene_create.argtypes = [P, P, P, P, C.c_uint32, P, P, C.c_uint32, P]
        lib.ref_reflect.argtypes = [P, P, P]
        lib.ref_scene_load.argtypes = [C.c_char_p, C.c_char_p]
        lib.ref_scene_load.restype = P
        lib.ref_custom_refract.argtypes = [P, P, C.c_float, P]
        lib.ref_scene_create.restype = P
        lib.ref_scene_free.argtypes = [P]
        lib.ref_scene_trace.argtypes = [P, P, C.c_uint32, P, P, P, P]
        lib.ref_scene_occluded.argtypes = [P, P, C.c_uint32, P]
        lib.ref_render_bdpt.argtypes = [P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_int, P]
        lib.ref_camera.argtypes = [P, C.c_float, C.c_float, C.c_uint32, C.c_uint32, P, P, P]
        lib.ref_load_scene.argtypes = [C.c_char_p, C.c_char_p, P, P, P, P, P, P]
        lib.ref_stl_load.argtypes = [C.c_char_p, C.c_char_p, P, P, P, P]
        lib.ref_stl_load.restype = C.c_int
        lib.ref_load_scene.restype = C.c_int
        lib.ref_render.argtypes = [P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P]
        _lib = lib
    return _lib


def intersect(v0, v1, v2, o, d):
    a = [np.ascontiguousarray(x, np.float32) for x in (v0, v1, v2, o, d)]
    p = np.zeros(3, np.float32)
    b = np.zeros(3, np.float32)
    hit = load().ref_intersect(*[x.ctypes.data for x in a], p.ctypes.data, b.ctypes.data)
    return bool(hit), p, b


class RefScene:
    """The reference's Scene/GraphicsObject/Mesh built from world-space arrays
    (one object per material, identity transform)."""

    def __init__(self, scene, weights=None):
        """weights: optional (M, 4) diffuse, reflect, transparent weights and
        refraction index per material (default diffuse only)."""
        self.keep = [np.ascontiguousarray(a) for a in (scene.pos, scene.nrm, scene.idx, scene.tri_mat)]
        self.colors = np.ascontiguousarray(scene.materials["color"].reshape(-1), np.float32)
        self.lights = np.ascontiguousarray(scene.materials["light_source"], np.uint32)
        self.weights = None if weights is None else np.ascontiguousarray(weights, np.float32).reshape(-1, 4)
        pos, nrm, idx, tm = self.keep
        self.h = load().ref_scene_create(pos.ctypes.data, nrm.ctypes.data, idx.ctypes.data, tm.ctypes.data,
                                         idx.shape[0], self.colors.ctypes.data, self.lights.ctypes.data,
                                         self.lights.shape[0],
                                         None if self.weights is None else self.weights.ctypes.data)

    @classmethod
    def from_file(cls, scene_path: str, basepath: str) -> "RefScene":
        """The scene as MeshManager reads it (meshes from basepath + ../res/obj/)."""
        self = cls.__new__(cls)
        self.keep = []
        self.h = load().ref_scene_load(str(scene_path).encode(), str(basepath).encode())
        return self

    def trace(self, rays: np.ndarray):
        n = rays.shape[0]
        hit = np.zeros(n, np.int32)
        p = np.zeros((n, 3), np.float32)
        nr = np.zeros((n, 3), np.float32)
        obj = np.zeros(n, np.int32)
        rays = np.ascontiguousarray(rays)
        load().ref_scene_trace(self.h, rays.ctypes.data, n, hit.ctypes.data, p.ctypes.data, nr.ctypes.data,
                               obj.ctypes.data)
        return hit.astype(bool), p, nr, obj

    def occluded(self, segs: np.ndarray) -> np.ndarray:
        """Scene::isOccluded per (n, 6) segment (start, end)."""
        segs = np.ascontiguousarray(segs, np.float32)
        out = np.zeros(segs.shape[0], np.uint8)
        load().ref_scene_occluded(self.h, segs.ctypes.data, segs.shape[0], out.ctypes.data)
        return out

    def render(self, cam, spp: int, depth: int, threads: int = 8) -> np.ndarray:
        vi = np.array(cam.view_inv[:], np.float32)
        pi = np.array(cam.proj_inv[:], np.float32)
        og = np.array(cam.origin[:], np.float32)
        rgb = np.zeros((cam.height, cam.width, 3), np.float32)
        load().ref_render(self.h, vi.ctypes.data, pi.ctypes.data, og.ctypes.data, cam.width, cam.height, spp,
                          depth, threads, rgb.ctypes.data)
        return rgb

    def render_bdpt(self, cam, spp: int, vision: int, light: int, max_depth: int, threads: int = 8) -> np.ndarray:
        vi = np.array(cam.view_inv[:], np.float32)
        pi = np.array(cam.proj_inv[:], np.float32)
        og = np.array(cam.origin[:], np.float32)
        rgb = np.zeros((cam.height, cam.width, 3), np.float32)
        load().ref_render_bdpt(self.h, vi.ctypes.data, pi.ctypes.data, og.ctypes.data, cam.width, cam.height, spp,
                               vision, light, max_depth, threads, rgb.ctypes.data)
        return rgb

    def __del__(self):
        try:
            load().ref_scene_free(self.h)
        except Exception:
            pass


def reflect(I, N) -> np.ndarray:
    i, n = (np.ascontiguousarray(x, np.float32) for x in (I, N))
    out = np.zeros(3, np.float32)
    load().ref_reflect(i.ctypes.data, n.ctypes.data, out.ctypes.data)
    return out


def custom_refract(I, N, eta: float) -> np.ndarray:
    i, n = (np.ascontiguousarray(x, np.float32) for x in (I, N))
    out = np.zeros(3, np.float32)
    load().ref_custom_refract(i.ctypes.data, n.ctypes.data, eta, out.ctypes.data)
    return out


def camera(position, angle0: float, angle1: float, w: int, h: int):
    vi = np.zeros(16, np.float32)
    pi = np.zeros(16, np.float32)
    og = np.zeros(3, np.float32)
    load().ref_camera(np.asarray(position, np.float32).ctypes.data, angle0, angle1, w, h, vi.ctypes.data,
                      pi.ctypes.data, og.ctypes.data)
    return vi, pi, og


def load_scene(scene_path: str, basepath: str):
    """World-space flattening through MeshManager + GraphicsObject::init."""
    lib = load()
    cnt = np.zeros(3, np.uint32)
    lib.ref_load_scene(scene_path.encode(), basepath.encode(), None, None, None, None, None, cnt.ctypes.data)
    V, T, O = (int(x) for x in cnt)
    pos = np.zeros((V, 3), np.float32)
    nrm = np.zeros((V, 3), np.float32)
    idx = np.zeros((T, 3), np.uint32)
    tri_obj = np.zeros(T, np.uint32)
    colors = np.zeros((O, 3), np.float32)
    lib.ref_load_scene(scene_path.encode(), basepath.encode(), pos.ctypes.data, nrm.ctypes.data, idx.ctypes.data,
                       tri_obj.ctypes.data, colors.ctypes.data, cnt.ctypes.data)
    return pos, nrm, idx, tri_obj, colors


def stl_load(workdir: str, name: str):
    """STLLoader on <workdir>/../res/stl/<name>: (pos (V, 3), nrm (V, 3),
    idx (T, 3)) in the loader's vertex order."""
    lib = load()
    cnt = np.zeros(2, np.uint32)
    rc = lib.ref_stl_load(str(workdir).encode(), name.encode(), None, None, None, cnt.ctypes.data)
    assert rc == 0, rc
    V, T = (int(x) for x in cnt)
    pos = np.zeros((V, 3), np.float32)
    nrm = np.zeros((V, 3), np.float32)
    idx = np.zeros((T, 3), np.uint32)
    rc = lib.ref_stl_load(str(workdir).encode(), name.encode(), pos.ctypes.data, nrm.ctypes.data, idx.ctypes.data,
                          cnt.ctypes.data)
    assert rc == 0, rc
    return pos, nrm, idx
