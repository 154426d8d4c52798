"""ctypes binding of include/rt.h (libmtrt.so).

The library is loaded from this package directory only.  There is no fallback:
if libmtrt.so has not been built, importing the binding raises, so nothing can
silently run on another code path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
# MTRT_LIB: another build of the library (for comparing kernel variants)
LIB_PATH = Path(os.environ["MTRT_LIB"]) if os.environ.get("MTRT_LIB") else PKG / "libmtrt.so"

RT_OK = 0
RT_E_INVALID = -1
RT_E_DEVICE = -2
RT_E_NOMEM = -3
RT_E_STATE = -4
RT_E_UNSUPPORTED = -5
RT_PTR_HOST = 0
RT_PTR_DEVICE = 1
RT_NOALLOC_SCRATCH_BYTES = 32 << 20  # rt_trace_closest_noalloc's fixed scratch per GPU
RT_MAX_RAYS = 0xFFFFFF00  # rays (or segments) per trace call
RT_MISS = 0xFFFFFFFF
RT_CHILD_EMPTY = 0xFFFFFFFF
RT_LEAF_BIT = 0x80000000
RT_LEAF_COUNT_SHIFT = 27
RT_LEAF_FIRST_MASK = 0x07FFFFFF
RT_RENDER_DETERMINISTIC = 0x1
RT_RENDER_SKY = 0x2
RT_RENDER_SUN = 0x4

_ERR_NAMES = {
    RT_E_INVALID: "RT_E_INVALID",
    RT_E_DEVICE: "RT_E_DEVICE",
    RT_E_NOMEM: "RT_E_NOMEM",
    RT_E_STATE: "RT_E_STATE",
    RT_E_UNSUPPORTED: "RT_E_UNSUPPORTED",
}

# numpy mirrors of the C structs (layouts asserted against ctypes below)
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("tri", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
MATERIAL_DTYPE = np.dtype([
    ("color", "<f4", 3), ("diffuse_threshold", "<f4"), ("reflect_threshold", "<f4"),
    ("transparent_threshold", "<f4"), ("refraction_index", "<f4"), ("light_source", "<u4"),
])
NODE_DTYPE = np.dtype([
    ("lo_x", "<f4", 4), ("hi_x", "<f4", 4), ("lo_y", "<f4", 4), ("hi_y", "<f4", 4),
    ("lo_z", "<f4", 4), ("hi_z", "<f4", 4), ("child", "<u4", 4), ("meta", "<u4", 4),
])
QNODE_DTYPE = np.dtype([
    ("origin", "<f4", 3), ("exp", "i1", 3), ("pad", "u1"), ("lo_x", "u1", 4), ("hi_x", "u1", 4),
    ("lo_y", "u1", 4), ("hi_y", "u1", 4), ("lo_z", "u1", 4), ("hi_z", "u1", 4), ("parent", "<u4"),
    ("reserved", "<u4"), ("child", "<u4", 4),
])
TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("id", "<u4"), ("e0", "<f4", 3), ("pad0", "<u4"),
                      ("e1", "<f4", 3), ("pad1", "<u4")])


class rt_ray(C.Structure):
    _fields_ = [("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("tmin", C.c_float),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float), ("tmax", C.c_float)]


class rt_hit(C.Structure):
    _fields_ = [("tri", C.c_uint32), ("t", C.c_float), ("u", C.c_float), ("v", C.c_float)]


class rt_material(C.Structure):
    _fields_ = [("color", C.c_float * 3), ("diffuse_threshold", C.c_float), ("reflect_threshold", C.c_float),
                ("transparent_threshold", C.c_float), ("refraction_index", C.c_float),
                ("light_source", C.c_uint32)]


class rt_scene_desc(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("nrm", C.c_void_p), ("idx", C.c_void_p), ("tri_mat", C.c_void_p),
                ("mat", C.c_void_p), ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32),
                ("num_materials", C.c_uint32)]


class rt_bvh4_node(C.Structure):
    _fields_ = [("lo_x", C.c_float * 4), ("hi_x", C.c_float * 4), ("lo_y", C.c_float * 4),
                ("hi_y", C.c_float * 4), ("lo_z", C.c_float * 4), ("hi_z", C.c_float * 4),
                ("child", C.c_uint32 * 4), ("meta", C.c_uint32 * 4)]


class rt_tri_rec(C.Structure):
    _fields_ = [("v0", C.c_float * 3), ("id", C.c_uint32), ("e0", C.c_float * 3), ("pad0", C.c_uint32),
                ("e1", C.c_float * 3), ("pad1", C.c_uint32)]


RT_BUILD_CLUSTER = 1


class rt_build_opts(C.Structure):
    _fields_ = [("leaf_size", C.c_uint32), ("flags", C.c_uint32)]


class rt_bvh_info(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("num_leaves", C.c_uint32), ("depth", C.c_uint32),
                ("num_triangles", C.c_uint32), ("build_ms", C.c_float), ("leaf_size", C.c_uint32)]


class rt_trace_counters(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("hits", C.c_uint64)]


class rt_camera(C.Structure):
    _fields_ = [("view_inv", C.c_float * 16), ("proj_inv", C.c_float * 16), ("origin", C.c_float * 3),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class rt_settings(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("spp", C.c_uint32), ("seed", C.c_uint64), ("flags", C.c_uint32),
                ("sample_offset", C.c_uint32), ("sky", C.c_float * 3), ("num_views", C.c_uint32),
                ("sun_dir", C.c_float * 3), ("sun_cos", C.c_float), ("sun_radiance", C.c_float * 3)]


class rt_bdpt_settings(C.Structure):
    _fields_ = [("vision_jumps", C.c_uint32), ("light_jumps", C.c_uint32), ("max_depth", C.c_uint32),
                ("spp", C.c_uint32), ("seed", C.c_uint64), ("sample_offset", C.c_uint32), ("flags", C.c_uint32)]


class rt_timing(C.Structure):
    _fields_ = [("last_kernel_ms", C.c_float), ("last_call_ms", C.c_float), ("launches", C.c_uint32),
                ("iterations", C.c_uint32), ("rays_traced", C.c_uint64)]


assert C.sizeof(rt_ray) == 32 and RAY_DTYPE.itemsize == 32
assert C.sizeof(rt_hit) == 16 and HIT_DTYPE.itemsize == 16
assert C.sizeof(rt_material) == 32 and MATERIAL_DTYPE.itemsize == 32
assert C.sizeof(rt_bvh4_node) == 128 and NODE_DTYPE.itemsize == 128
assert C.sizeof(rt_tri_rec) == 48 and TRI_DTYPE.itemsize == 48
assert QNODE_DTYPE.itemsize == 64

class rt_nrc_config(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("seed", C.c_uint64), ("bounds_lo", C.c_float * 3), ("bounds_hi", C.c_float * 3)]


RT_NRC_FEATURES = 9
RT_NRC_NO_TRAIN = 0x1
RT_NRC_NO_QUERY = 0x2
RT_NRC_NO_RECORDS = 0x4
RT_NRC_NO_NEE = 0x8


class rt_nrc_render(C.Structure):
    _fields_ = [("query_vertex", C.c_uint32), ("train_every", C.c_uint32), ("batch", C.c_uint32),
                ("max_steps", C.c_uint32), ("flags", C.c_uint32)]


class rt_nrc_stats(C.Structure):
    _fields_ = [("queries", C.c_uint64), ("records", C.c_uint64), ("steps", C.c_uint32), ("loss", C.c_float),
                ("render_ms", C.c_float), ("train_ms", C.c_float), ("shadow_rays", C.c_uint64)]

class rt_ddgi_config(C.Structure):
    _fields_ = [("probe_count", C.c_uint32 * 3), ("probe_start", C.c_float * 3), ("probe_spacing", C.c_float * 3),
                ("rays_per_probe", C.c_uint32), ("side", C.c_uint32), ("max_ray_distance", C.c_float),
                ("depth_sharpness", C.c_float), ("normal_bias", C.c_float), ("energy_preservation", C.c_float),
                ("hysteresis", C.c_float), ("shadow_count", C.c_uint32), ("flags", C.c_uint32)]


RT_DDGI_LINEAR_BLENDING = 0x1
RT_DDGI_MULTI_BOUNCE = 0x2
RT_DDGI_ROTATE = 0x4
RT_DDGI_REFERENCE_VISIBILITY = 0x8
RT_DDGI_SURFEL_FLOATS = 8

# every exported symbol of include/rt.h with its signature
_P = C.c_void_p
SIGNATURES = {
    "rt_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "rt_destroy": (None, [_P]),
    "rt_last_error": (C.c_char_p, [_P]),
    "rt_set_stream": (C.c_int, [_P, _P]),
    "rt_get_timing": (C.c_int, [_P, C.POINTER(rt_timing)]),
    "rt_scene_upload": (C.c_int, [_P, C.POINTER(rt_scene_desc)]),
    "rt_set_material_colors": (C.c_int, [_P, _P, C.c_int]),
    "rt_build": (C.c_int, [_P, C.POINTER(rt_build_opts), C.POINTER(rt_bvh_info)]),
    "rt_bvh_export": (C.c_int, [_P, _P, C.POINTER(C.c_uint32), _P, C.POINTER(C.c_uint32)]),
    "rt_bvh_export_traversal": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "rt_trace_closest": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_int, C.POINTER(rt_trace_counters)]),
    "rt_trace_closest_noalloc": (C.c_int, [_P, _P, C.c_uint32, _P, _P]),
    "rt_trace_any": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_int]),
    "rt_occluded": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_int]),
    "rt_render": (C.c_int, [_P, C.POINTER(rt_camera), C.POINTER(rt_settings), _P, _P, C.c_int]),
    "rt_render_views": (C.c_int, [_P, C.POINTER(rt_camera), C.c_uint32, C.POINTER(rt_settings), _P, _P, C.c_int]),
    "rt_render_bdpt": (C.c_int, [_P, C.POINTER(rt_camera), C.POINTER(rt_bdpt_settings), _P, _P, C.c_int]),
    "rt_render_adjoint": (C.c_int, [_P, C.POINTER(rt_camera), C.POINTER(rt_settings), _P, _P, C.c_int]),
    "rt_nrc_create": (C.c_int, [_P, C.POINTER(rt_nrc_config), C.POINTER(_P)]),
    "rt_nrc_destroy": (None, [_P]),
    "rt_nrc_params": (C.c_int, [_P, _P, C.POINTER(C.c_uint32), C.c_int]),
    "rt_nrc_infer": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_int]),
    "rt_nrc_gradient": (C.c_int, [_P, _P, _P, C.c_uint32, _P, C.POINTER(C.c_float), C.c_int]),
    "rt_nrc_train": (C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(C.c_float), C.c_int]),
    "rt_render_nrc_records": (C.c_int, [_P, _P, _P, C.POINTER(C.c_uint32)]),
    "rt_render_nrc": (C.c_int, [_P, _P, C.POINTER(rt_camera), C.c_uint32, C.POINTER(rt_settings),
                                C.POINTER(rt_nrc_render), _P, _P, C.c_int, C.POINTER(rt_nrc_stats)]),
    "rt_denoise_gauss_weights": (C.c_int, [C.c_uint32, C.c_float, _P]),
    "rt_denoise_gauss": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_int]),
    "rt_denoise_median": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_uint32, C.c_int]),
    "rt_ddgi_create": (C.c_int, [_P, C.POINTER(rt_ddgi_config), C.POINTER(_P)]),
    "rt_ddgi_destroy": (None, [_P]),
    "rt_ddgi_atlas_size": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "rt_ddgi_update": (C.c_int, [_P, C.POINTER(rt_settings)]),
    "rt_ddgi_update_surfels": (C.c_int, [_P, _P, C.c_int]),
    "rt_ddgi_surfels": (C.c_int, [_P, _P, _P, C.c_int]),
    "rt_ddgi_get_atlas": (C.c_int, [_P, _P, _P, C.c_int]),
    "rt_ddgi_set_atlas": (C.c_int, [_P, _P, _P, C.c_int]),
    "rt_ddgi_eval": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_int]),
    "rt_render_ddgi": (C.c_int, [_P, _P, C.POINTER(rt_camera), C.c_uint32, C.POINTER(rt_settings), C.c_uint32,
                                 _P, _P, C.c_int, C.POINTER(rt_nrc_stats)]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libmtrt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        # torch ships its own HIP runtime under the same SONAME
        # (libamdhip64.so.7).  Loading torch first makes libmtrt bind to that
        # one copy; loaded the other way round the process ends up with two
        # runtimes and torch cannot open the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class RTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


def _ptr(a) -> int:
    """Address of a numpy array (C-contiguous) or a torch tensor (contiguous)."""
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
        return a.ctypes.data
    assert a.is_contiguous(), "tensor must be contiguous"
    return a.data_ptr()


def _is_device(a) -> bool:
    return not isinstance(a, np.ndarray) and getattr(a, "is_cuda", False)
