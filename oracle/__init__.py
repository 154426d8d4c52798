"""ctypes loader of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libref.so"

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        lib = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        F = C.POINTER(C.c_float)
        U = C.POINTER(C.c_uint32)
        lib.or_philox.argtypes = [U, U, U]
        lib.or_sincos_2pi.argtypes = [C.c_float, F, F]
        lib.or_hemisphere.argtypes = [C.c_float, C.c_float, F, F]
        lib.or_intersect.argtypes = [F, F, F, F, F, F]
        lib.or_intersect.restype = C.c_int
        lib.or_barycentric.argtypes = [F, F, F, F, F]
        lib.or_reflect.argtypes = [F, F, F]
        lib.or_trace_brute.argtypes = [P, P, C.c_uint32, P, C.c_uint32, P, C.c_int]
        lib.or_bvh_build.argtypes = [P, P, C.c_uint32, C.c_int]
        lib.or_bvh_build.restype = P
        lib.or_bvh_free.argtypes = [P]
        lib.or_bvh_num_nodes.argtypes = [P]
        lib.or_bvh_num_nodes.restype = C.c_uint32
        lib.or_trace_bvh.argtypes = [P, P, C.c_uint32, P, C.c_int, C.POINTER(C.c_uint64)]
        lib.or_trace_bvh4.argtypes = [P, C.c_uint32, P, C.c_uint32, C.c_uint32, P, C.c_uint32, P, C.c_int,
                                      C.POINTER(C.c_uint64)]
        lib.or_check_bvh4.argtypes = [P, C.c_uint32, P, C.c_uint32, P, P, U]
        lib.or_check_bvh4.restype = C.c_int
        lib.or_vec4_mat.argtypes = [F, F, F]
        lib.or_primary_dir.argtypes = [F, F, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, F]
        lib.or_render.argtypes = [P, P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P, P, C.c_int]
        lib.or_render.restype = C.c_int
        lib.or_render_adjoint.argtypes = [P, P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P, P, C.c_int]
        lib.or_render_adjoint.restype = C.c_int
        _lib = lib
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    load().or_philox(c.ctypes.data_as(C.POINTER(C.c_uint32)), k.ctypes.data_as(C.POINTER(C.c_uint32)),
                     out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def sincos_2pi(x: float):
    s, c = C.c_float(), C.c_float()
    load().or_sincos_2pi(x, C.byref(s), C.byref(c))
    return s.value, c.value


def intersect(o, d, v0, v1, v2):
    arrs = [np.asarray(a, np.float32) for a in (o, d, v0, v1, v2)]
    out = np.zeros(3, np.float32)
    hit = load().or_intersect(*[fptr(a) for a in arrs], fptr(out))
    return bool(hit), out


def reflect(I, N) -> np.ndarray:
    i, n = (np.asarray(a, np.float32) for a in (I, N))
    out = np.zeros(3, np.float32)
    load().or_reflect(fptr(i), fptr(n), fptr(out))
    return out


def custom_refract(I, N, eta: float) -> np.ndarray:
    lib = load()
    lib.or_custom_refract.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.POINTER(C.c_float)]
    i, n = (np.asarray(a, np.float32) for a in (I, N))
    out = np.zeros(3, np.float32)
    lib.or_custom_refract(fptr(i), fptr(n), eta, fptr(out))
    return out


def sin_acos(x: float) -> float:
    lib = load()
    lib.or_sin_acos.argtypes = [C.c_double]
    lib.or_sin_acos.restype = C.c_double
    return lib.or_sin_acos(x)


def barycentric(v0, v1, v2, p):
    arrs = [np.asarray(a, np.float32) for a in (v0, v1, v2, p)]
    out = np.zeros(3, np.float32)
    load().or_barycentric(*[fptr(a) for a in arrs], fptr(out))
    return out


def hit_shading(scene, rays: np.ndarray, hits: np.ndarray):
    """(hit points, shading normals) of closest hits, as Scene::traceRay forms them."""
    lib = load()
    lib.or_hit_shading.argtypes = [C.c_void_p] * 5 + [C.c_uint32, C.c_void_p, C.c_void_p]
    n = rays.shape[0]
    p = np.zeros((n, 3), np.float32)
    nr = np.zeros((n, 3), np.float32)
    arrs = [np.ascontiguousarray(a) for a in (scene.pos, scene.nrm, scene.idx, rays, hits)]
    lib.or_hit_shading(*[_p(a) for a in arrs], n, _p(p), _p(nr))
    return p, nr


def trace_brute(scene, rays: np.ndarray, threads: int = 8) -> np.ndarray:
    from master_thesis_amd._lib import HIT_DTYPE
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    pos = np.ascontiguousarray(scene.pos, np.float32)
    idx = np.ascontiguousarray(scene.idx, np.uint32)
    load().or_trace_brute(_p(pos), _p(idx), idx.shape[0], _p(np.ascontiguousarray(rays)), rays.shape[0], _p(hits),
                          threads)
    return hits


def occluded_brute(scene, segs: np.ndarray, threads: int = 8) -> np.ndarray:
    """Scene::isOccluded by brute force for (n, 6) segments (start, end)."""
    lib = load()
    P = C.c_void_p
    lib.or_occluded_brute.argtypes = [P, P, C.c_uint32, P, C.c_uint32, P, C.c_int]
    lib.or_occluded_brute.restype = None
    segs = np.ascontiguousarray(segs, np.float32)
    out = np.zeros(segs.shape[0], np.uint8)
    pos = np.ascontiguousarray(scene.pos, np.float32)
    idx = np.ascontiguousarray(scene.idx, np.uint32)
    lib.or_occluded_brute(_p(pos), _p(idx), idx.shape[0], _p(segs), segs.shape[0], _p(out), threads)
    return out


class CpuBVH:
    """The oracle's binned-SAH BVH (the CPU baseline's acceleration structure)."""

    def __init__(self, scene, leaf_size: int = 4):
        self.pos = np.ascontiguousarray(scene.pos, np.float32)
        self.idx = np.ascontiguousarray(scene.idx, np.uint32)
        self.h = load().or_bvh_build(_p(self.pos), _p(self.idx), self.idx.shape[0], leaf_size)

    def trace(self, rays: np.ndarray, threads: int = 8):
        from master_thesis_amd._lib import HIT_DTYPE
        hits = np.zeros(rays.shape[0], HIT_DTYPE)
        cnt = (C.c_uint64 * 4)()
        load().or_trace_bvh(self.h, _p(np.ascontiguousarray(rays)), rays.shape[0], _p(hits), threads, cnt)
        return hits, dict(rays=cnt[0], node_visits=cnt[1], tri_tests=cnt[2], hits=cnt[3])

    def __del__(self):
        try:
            load().or_bvh_free(self.h)
        except Exception:
            pass


def trace_bvh4(nodes, tris, depth: int, rays: np.ndarray, threads: int = 8):
    """Replay of the GPU walk over exported traversal nodes (QNODE_DTYPE)."""
    from master_thesis_amd._lib import HIT_DTYPE
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    cnt = (C.c_uint64 * 4)()
    load().or_trace_bvh4(_p(nodes), nodes.shape[0], _p(tris), tris.shape[0], depth, _p(np.ascontiguousarray(rays)),
                         rays.shape[0], _p(hits), threads, cnt)
    return hits, dict(rays=cnt[0], node_visits=cnt[1], tri_tests=cnt[2], hits=cnt[3])


def check_quantized(qnodes, nodes) -> int:
    """0 when the traversal's quantised nodes stand soundly for the build nodes."""
    lib = load()
    lib.or_check_quantized.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    lib.or_check_quantized.restype = C.c_int
    assert qnodes.shape[0] == nodes.shape[0]
    return lib.or_check_quantized(_p(qnodes), _p(nodes), nodes.shape[0])


def check_bvh4(nodes, tris, scene):
    d = C.c_uint32(0)
    pos = np.ascontiguousarray(scene.pos, np.float32)
    idx = np.ascontiguousarray(scene.idx, np.uint32)
    rc = load().or_check_bvh4(_p(nodes), nodes.shape[0], _p(tris), tris.shape[0], _p(pos), _p(idx), C.byref(d))
    return rc, d.value


def primary_dir(cam, px: int, py: int) -> np.ndarray:
    vi = np.array(cam.view_inv[:], np.float32)
    pi = np.array(cam.proj_inv[:], np.float32)
    out = np.zeros(3, np.float32)
    load().or_primary_dir(fptr(vi), fptr(pi), px, py, cam.width, cam.height, fptr(out))
    return out


def _scene_ptrs(scene):
    arrs = [np.ascontiguousarray(a) for a in (scene.pos, scene.nrm, scene.idx, scene.tri_mat, scene.materials)]
    return arrs, [_p(a) for a in arrs]


def render(scene, cam, settings, threads: int = 8):
    """Returns (rgb float32 HxWx3, rgb8 uint8 HxWx3)."""
    arrs, ps = _scene_ptrs(scene)
    H, W = cam.height, cam.width
    rgb = np.zeros((H, W, 3), np.float32)
    rgb8 = np.zeros((H, W, 3), np.uint8)
    rc = load().or_render(*ps, arrs[0].shape[0], arrs[2].shape[0], arrs[4].shape[0], C.addressof(cam),
                          C.addressof(settings), _p(rgb), _p(rgb8), threads)
    assert rc == 0, rc
    return rgb, rgb8


def render_views(scene, cams, settings, threads: int = 8):
    """Returns (rgb (V, H, W, 3) float32, rgb8 uint8) for a list of cameras."""
    lib = load()
    P = C.c_void_p
    lib.or_render_views.argtypes = [P, P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, P, C.c_uint32, P, P, P,
                                    C.c_int]
    lib.or_render_views.restype = C.c_int
    arrs, ps = _scene_ptrs(scene)
    from master_thesis_amd._lib import rt_camera
    V = len(cams)
    carr = (rt_camera * V)(*cams)
    H, W = cams[0].height, cams[0].width
    rgb = np.zeros((V, H, W, 3), np.float32)
    rgb8 = np.zeros((V, H, W, 3), np.uint8)
    rc = lib.or_render_views(*ps, arrs[0].shape[0], arrs[2].shape[0], arrs[4].shape[0], C.addressof(carr), V,
                             C.addressof(settings), _p(rgb), _p(rgb8), threads)
    assert rc == 0, rc
    return rgb, rgb8


def last_render_seconds() -> float:
    """Wall time of the last render / render_views pass, BVH build excluded."""
    lib = load()
    lib.or_last_render_seconds.restype = C.c_double
    lib.or_last_render_seconds.argtypes = []
    return float(lib.or_last_render_seconds())


def render_adjoint(scene, cam, settings, d_image: np.ndarray, threads: int = 8) -> np.ndarray:
    arrs, ps = _scene_ptrs(scene)
    M = arrs[4].shape[0]
    d_color = np.zeros((M, 3), np.float32)
    d_image = np.ascontiguousarray(d_image, np.float32)
    rc = load().or_render_adjoint(*ps, arrs[0].shape[0], arrs[2].shape[0], M, C.addressof(cam),
                                  C.addressof(settings), _p(d_image), _p(d_color), threads)
    assert rc == 0, rc
    return d_color


def render_cached(scene, cams, settings, cache_rgb, q_vertex: int = 1, train_every: int = 16, query: bool = True,
                  nee: bool = False, records: bool = False, threads: int = 8):
    """The cache mode of rt_render_nrc with a constant cache prediction.
    Returns (rgb (V, H, W, 3) float32, queries, records) and, with
    records=True, also (feats (n, 9), targets (n, 3), keys (n,)) of the
    training records in (view, pixel, sample) order, key = (view * H * W +
    pixel) * spp + sample."""
    lib = load()
    P = C.c_void_p
    lib.or_render_cached.argtypes = [P, P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, P, C.c_uint32, P,
                                     C.c_uint32, C.c_uint32, C.c_int, C.c_int, P, P, P, P, C.c_int]
    lib.or_render_cached.restype = C.c_int
    arrs, ps = _scene_ptrs(scene)
    from master_thesis_amd._lib import rt_camera
    if not isinstance(cams, (list, tuple)):
        cams = [cams]
    V = len(cams)
    carr = (rt_camera * V)(*cams)
    H, W = cams[0].height, cams[0].width
    rgb = np.zeros((V, H, W, 3), np.float32)
    c = np.ascontiguousarray(cache_rgb, np.float32)
    counts = np.zeros(2, np.uint64)
    rec = np.zeros((V * H * W * settings.spp, 13), np.float32) if records else None
    rc = lib.or_render_cached(*ps, arrs[0].shape[0], arrs[2].shape[0], arrs[4].shape[0], C.addressof(carr), V,
                              C.addressof(settings), q_vertex, train_every, int(query), int(nee), _p(c), _p(rgb),
                              _p(rec) if records else None, _p(counts), threads)
    assert rc == 0, rc
    if records:
        ok = rec[:, 12] == 1.0
        return rgb, int(counts[0]), int(counts[1]), rec[ok, :9], rec[ok, 9:12], np.nonzero(ok)[0]
    return rgb, int(counts[0]), int(counts[1])


def render_faithful(scene, cam, spp: int, max_depth: int, threads: int = 8, seed: int = 1,
                    pixel_limit: int | None = None):
    """The reference renderer in faithful mode (oracle/faithful.cpp): its
    two-level agglomerative BVHs, enumerate-all queries with a vector per
    node, one mutex-guarded mt19937, one pixel per atomic fetch_add on
    `threads` workers.  Objects as the scene groups them (Scene.objects()).
    Renders the first pixel_limit pixels (default all).  Returns (rgb (H, W,
    3) float32, rgb8 (H, W, 3) uint8, seconds without the BVH builds)."""
    lib = load()
    P = C.c_void_p
    lib.or_render_faithful.argtypes = [P, P, P, P, C.c_uint32, P, C.c_uint32, P, P, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_uint32, P, P]
    lib.or_render_faithful.restype = C.c_int
    lib.or_faithful_seconds.restype = C.c_double
    tri_obj, obj_mat = scene.objects()
    arrs = [np.ascontiguousarray(a) for a in (scene.pos, scene.nrm, scene.idx, tri_obj.astype(np.uint32),
                                              obj_mat.astype(np.uint32), scene.materials)]
    H, W = cam.height, cam.width
    rgb = np.zeros((H, W, 3), np.float32)
    rgb8 = np.zeros((H, W, 3), np.uint8)
    lim = W * H if pixel_limit is None else int(pixel_limit)
    rc = lib.or_render_faithful(_p(arrs[0]), _p(arrs[1]), _p(arrs[2]), _p(arrs[3]), arrs[2].shape[0], _p(arrs[4]),
                                arrs[4].shape[0], _p(arrs[5]), C.addressof(cam), spp, max_depth, threads, seed, lim,
                                _p(rgb), _p(rgb8))
    assert rc == 0, rc
    return rgb, rgb8, float(lib.or_faithful_seconds())


def render_bdpt(scene, cam, settings, vision_jumps: int = 5, light_jumps: int = 5, max_depth: int = 5,
                threads: int = 8) -> np.ndarray:
    """The bidirectional path tracer (rt_render_bdpt's checker); rgb (H, W, 3)."""
    lib = load()
    P = C.c_void_p
    lib.or_render_bdpt.argtypes = [P, P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, C.c_uint32, C.c_uint32,
                                   C.c_uint32, P, C.c_int]
    lib.or_render_bdpt.restype = C.c_int
    arrs, ps = _scene_ptrs(scene)
    rgb = np.zeros((cam.height, cam.width, 3), np.float32)
    rc = lib.or_render_bdpt(*ps, arrs[0].shape[0], arrs[2].shape[0], arrs[4].shape[0], C.addressof(cam),
                            C.addressof(settings), vision_jumps, light_jumps, max_depth, _p(rgb), threads)
    assert rc == 0, rc
    return rgb



# ---- irradiance probes (rt_ddgi_*): cfg is a master_thesis_amd._lib.rt_ddgi_config

def ddgi_atlas_shape(cfg):
    ts = cfg.side + 2
    return cfg.probe_count[2] * ts, cfg.probe_count[0] * cfg.probe_count[1] * ts


def ddgi_fib(i: int, n: int) -> np.ndarray:
    lib = load()
    lib.or_ddgi_fib.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p]
    out = np.zeros(3, np.float32)
    lib.or_ddgi_fib(i, n, _p(out))
    return out


def ddgi_oct_encode(v) -> np.ndarray:
    lib = load()
    lib.or_ddgi_oct_encode.argtypes = [C.c_void_p, C.c_void_p]
    v = np.ascontiguousarray(v, np.float32)
    o = np.zeros(2, np.float32)
    lib.or_ddgi_oct_encode(_p(v), _p(o))
    return o


def ddgi_oct_decode(o) -> np.ndarray:
    lib = load()
    lib.or_ddgi_oct_decode.argtypes = [C.c_void_p, C.c_void_p]
    o = np.ascontiguousarray(o, np.float32)
    v = np.zeros(3, np.float32)
    lib.or_ddgi_oct_decode(_p(o), _p(v))
    return v


def ddgi_update(cfg, surfels: np.ndarray, irr=None, dep=None):
    """Atlases (irr (H, W, 4), dep (H, W, 2)) after an update from surfels
    (probes * rays, 8); irr / dep: the previous atlases (None: first update)."""
    lib = load()
    lib.or_ddgi_update.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    H, W = ddgi_atlas_shape(cfg)
    first = irr is None
    irr = np.zeros((H, W, 4), np.float32) if first else np.array(irr, np.float32, copy=True)
    dep = np.zeros((H, W, 2), np.float32) if first else np.array(dep, np.float32, copy=True)
    s = np.ascontiguousarray(surfels, np.float32)
    lib.or_ddgi_update(C.addressof(cfg), _p(s), int(first), _p(irr), _p(dep))
    return irr, dep


def ddgi_eval(cfg, irr, dep, samples) -> np.ndarray:
    lib = load()
    lib.or_ddgi_eval.argtypes = [C.c_void_p] * 4 + [C.c_uint32, C.c_void_p]
    irr = np.ascontiguousarray(irr, np.float32)
    dep = np.ascontiguousarray(dep, np.float32)
    s = np.ascontiguousarray(samples, np.float32).reshape(-1, 9)
    out = np.zeros((s.shape[0], 3), np.float32)
    lib.or_ddgi_eval(C.addressof(cfg), _p(irr), _p(dep), _p(s), s.shape[0], _p(out))
    return out


def ddgi_surfels(scene, cfg, settings, rot, update: int = 0, irr=None, dep=None, multi: bool = False,
                 threads: int = 8) -> np.ndarray:
    """Surfels (probes * rays, 8) of one field update on the CPU."""
    lib = load()
    P = C.c_void_p
    lib.or_ddgi_surfels.argtypes = [P, P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P, C.c_uint32, P, P,
                                    C.c_int, P, C.c_int]
    lib.or_ddgi_surfels.restype = C.c_int
    arrs, ps = _scene_ptrs(scene)
    n = cfg.probe_count[0] * cfg.probe_count[1] * cfg.probe_count[2] * cfg.rays_per_probe
    out = np.zeros((n, 8), np.float32)
    r = np.ascontiguousarray(rot, np.float32).reshape(9)
    H, W = ddgi_atlas_shape(cfg)
    irr = np.zeros((H, W, 4), np.float32) if irr is None else np.ascontiguousarray(irr, np.float32)
    dep = np.zeros((H, W, 2), np.float32) if dep is None else np.ascontiguousarray(dep, np.float32)
    rc = lib.or_ddgi_surfels(*ps, arrs[0].shape[0], arrs[2].shape[0], arrs[4].shape[0], C.addressof(cfg),
                             C.addressof(settings), _p(r), update, _p(irr), _p(dep), int(multi), _p(out), threads)
    assert rc == 0, rc
    return out


def render_ddgi(scene, cams, settings, cfg, irr, dep, q_vertex: int = 0, threads: int = 8):
    """rt_render_ddgi's checker: (rgb (V, H, W, 3), queries)."""
    lib = load()
    P = C.c_void_p
    lib.or_render_ddgi.argtypes = [P, P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, P, C.c_uint32, P, C.c_uint32,
                                   P, P, P, P, P, C.c_int]
    lib.or_render_ddgi.restype = C.c_int
    arrs, ps = _scene_ptrs(scene)
    from master_thesis_amd._lib import rt_camera
    if not isinstance(cams, (list, tuple)):
        cams = [cams]
    V = len(cams)
    carr = (rt_camera * V)(*cams)
    H, W = cams[0].height, cams[0].width
    rgb = np.zeros((V, H, W, 3), np.float32)
    q = np.zeros(1, np.uint64)
    irr = np.ascontiguousarray(irr, np.float32)
    dep = np.ascontiguousarray(dep, np.float32)
    rc = lib.or_render_ddgi(*ps, arrs[0].shape[0], arrs[2].shape[0], arrs[4].shape[0], C.addressof(carr), V,
                            C.addressof(settings), q_vertex, C.addressof(cfg), _p(irr), _p(dep), _p(rgb), _p(q),
                            threads)
    assert rc == 0, rc
    return rgb, int(q[0])


# ---- post-filters

def gauss_weights(k: int, sigma: float) -> np.ndarray:
    lib = load()
    lib.or_gauss_weights.argtypes = [C.c_uint32, C.c_float, C.c_void_p]
    g = np.zeros(20, np.float32)
    assert lib.or_gauss_weights(k, sigma, _p(g)) == 0
    return g[:k]


def denoise_gauss(img: np.ndarray, k: int, sigma: float) -> np.ndarray:
    lib = load()
    lib.or_denoise_gauss.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float]
    a = np.ascontiguousarray(img, np.float32)
    out = np.empty_like(a)
    assert lib.or_denoise_gauss(_p(a), _p(out), a.shape[1], a.shape[0], k, sigma) == 0
    return out


def denoise_median(img: np.ndarray) -> np.ndarray:
    lib = load()
    lib.or_denoise_median.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]
    a = np.ascontiguousarray(img, np.float32)
    out = np.empty_like(a)
    assert lib.or_denoise_median(_p(a), _p(out), a.shape[1], a.shape[0]) == 0
    return out
