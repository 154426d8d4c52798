"""ctypes binding of include/rt_host.h (libmtrt_host.so): scene files,
cameras and .renderer files in the reference's formats."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _lib as L
from .scenes import Scene

LIB_PATH = Path(__file__).resolve().parent / "libmtrt_host.so"
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m master_thesis_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        lib.mth_scene_load.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(P)]
        lib.mth_scene_load.restype = C.c_int
        lib.mth_mesh_load.argtypes = [C.c_char_p, C.POINTER(P)]
        lib.mth_mesh_load.restype = C.c_int
        lib.mth_scene_desc.argtypes = [P, C.POINTER(L.rt_scene_desc)]
        lib.mth_scene_desc.restype = C.c_int
        lib.mth_scene_free.argtypes = [P]
        lib.mth_scene_free.restype = None
        lib.mth_camera_load.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(L.rt_camera)]
        lib.mth_camera_load.restype = C.c_int
        lib.mth_camera_make.argtypes = [C.POINTER(C.c_float), C.c_float, C.c_float, C.c_float, C.c_uint32,
                                        C.c_uint32, C.POINTER(L.rt_camera)]
        lib.mth_camera_make.restype = C.c_int
        lib.mth_renderer_load.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(L.rt_settings)]
        lib.mth_renderer_load.restype = C.c_int
        lib.mth_write_ppm.argtypes = [C.c_char_p, P, C.c_uint32, C.c_uint32]
        lib.mth_write_ppm.restype = C.c_int
        lib.mth_last_error.argtypes = []
        lib.mth_last_error.restype = C.c_char_p
        _lib = lib
    return _lib


def _check(rc: int) -> None:
    if rc != L.RT_OK:
        raise L.RTError(rc, load().mth_last_error().decode())


def load_scene(scene_path: str, mesh_dir: str) -> Scene:
    lib = load()
    h = C.c_void_p()
    _check(lib.mth_scene_load(str(scene_path).encode(), str(mesh_dir).encode(), C.byref(h)))
    return _take_scene(lib, h)


def load_mesh(mesh_path: str) -> Scene:
    """One OBJ or binary STL mesh in its own coordinates (mth_mesh_load)."""
    lib = load()
    h = C.c_void_p()
    _check(lib.mth_mesh_load(str(mesh_path).encode(), C.byref(h)))
    return _take_scene(lib, h)


def _take_scene(lib, h) -> Scene:
    try:
        d = L.rt_scene_desc()
        _check(lib.mth_scene_desc(h, C.byref(d)))
        V, T, M = d.num_vertices, d.num_triangles, d.num_materials

        def arr(ptr, n, dt):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(dt)), shape=(n,)).copy()

        pos = arr(d.pos, 3 * V, C.c_float).reshape(V, 3)
        nrm = arr(d.nrm, 3 * V, C.c_float).reshape(V, 3)
        idx = arr(d.idx, 3 * T, C.c_uint32).reshape(T, 3)
        tm = arr(d.tri_mat, T, C.c_uint32)
        mat = np.frombuffer(C.string_at(d.mat, M * L.MATERIAL_DTYPE.itemsize), L.MATERIAL_DTYPE).copy()
        s = Scene(pos=pos.astype(np.float32), nrm=nrm.astype(np.float32), idx=idx.astype(np.uint32),
                  tri_mat=tm.astype(np.uint32), materials=mat)
        s.validate()
        return s
    finally:
        lib.mth_scene_free(h)


def camera(position, angle0: float, angle1: float, fov: float, width: int, height: int) -> L.rt_camera:
    """Camera from values in the .camera file's terms (fov in radians)."""
    p = (C.c_float * 3)(*position)
    c = L.rt_camera()
    _check(load().mth_camera_make(p, angle0, angle1, fov, width, height, C.byref(c)))
    return c


def camera_from_file(path: str, width: int, height: int) -> L.rt_camera:
    c = L.rt_camera()
    _check(load().mth_camera_load(str(path).encode(), width, height, C.byref(c)))
    return c


def renderer_from_file(path: str):
    name = C.create_string_buffer(256)
    s = L.rt_settings()
    _check(load().mth_renderer_load(str(path).encode(), name, 256, C.byref(s)))
    return name.value.decode(), s


def write_ppm(path: str, rgb8: np.ndarray) -> None:
    rgb8 = np.ascontiguousarray(rgb8, np.uint8)
    H, W = rgb8.shape[:2]
    _check(load().mth_write_ppm(str(path).encode(), rgb8.ctypes.data, W, H))


def cornell_camera(width: int = 256, height: int = 256, fov_deg: float = 40.0) -> L.rt_camera:
    """The Cornell-style box's camera: on the +x axis looking -x (angle(1, 0)
    in the reference's camera terms), 40-degree field of view."""
    from .scenes import cornell_camera_position
    return camera(cornell_camera_position(), 1.0, 0.0, float(np.radians(fov_deg)), width, height)


def closed_camera(width: int = 256, height: int = 256, fov_deg: float = 75.0) -> L.rt_camera:
    """Camera inside the closed Cornell-style box, just in front of its
    closing wall, looking -x."""
    from .scenes import CORNELL_HALF
    return camera((CORNELL_HALF - 0.05, 0.0, 0.0), 1.0, 0.0, float(np.radians(fov_deg)), width, height)
