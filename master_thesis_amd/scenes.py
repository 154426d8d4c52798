"""Synthetic scenes and ray sets of the benchmark configurations.

Everything is generated here, seed-fixed; nothing is read from the reference.

  cornell_box()        Cornell-style box: 5 walls of a 5.5-unit cube (10 tris),
                       two axis-aligned boxes (24 tris), one emissive quad (2 tris,
                       colour 15), Lambertian albedos 0.73 / 0.65 / 0.12 triples.
  triangle_soup()      T triangles, centroids uniform in the unit cube, vertex
                       offsets of LogNormal(-5, 0.5) length in uniform directions.
  microbench_rays()    8M coherent pinhole rays + 8M uniform-sphere incoherent rays.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._lib import MATERIAL_DTYPE, RAY_DTYPE


@dataclass
class Scene:
    pos: np.ndarray        # (V, 3) float32, world space
    nrm: np.ndarray        # (V, 3) float32, world space
    idx: np.ndarray        # (T, 3) uint32
    tri_mat: np.ndarray    # (T,) uint32
    materials: np.ndarray  # (M,) MATERIAL_DTYPE
    # (T,) uint32 object of each triangle as the reference's scene files would
    # group them (one GraphicsObject per quad or box); None: one per material
    tri_obj: np.ndarray | None = None

    def objects(self) -> tuple[np.ndarray, np.ndarray]:
        """(object per triangle, material per object)."""
        if self.tri_obj is None:
            return self.tri_mat.copy(), np.arange(len(self.materials), dtype=np.uint32)
        O = int(self.tri_obj.max()) + 1
        obj_mat = np.zeros(O, np.uint32)
        obj_mat[self.tri_obj] = self.tri_mat
        return self.tri_obj, obj_mat

    @property
    def num_triangles(self) -> int:
        return int(self.idx.shape[0])

    def validate(self) -> None:
        assert self.pos.dtype == np.float32 and self.pos.ndim == 2 and self.pos.shape[1] == 3
        assert self.nrm.dtype == np.float32 and self.nrm.shape == self.pos.shape
        assert self.idx.dtype == np.uint32 and self.idx.ndim == 2 and self.idx.shape[1] == 3
        assert self.tri_mat.dtype == np.uint32 and self.tri_mat.shape == (self.idx.shape[0],)
        assert self.materials.dtype == MATERIAL_DTYPE


def material(color, diffuse=1.0, reflect=0.0, transparent=0.0, refraction_index=1.0, light=False) -> np.ndarray:
    """One material record with thresholds as GraphicsObject::init computes
    them (graphics_object.cpp:48-51), in float."""
    m = np.zeros((), MATERIAL_DTYPE)
    m["color"] = np.asarray(color, np.float32)
    wd, wr, wt = np.float32(diffuse), np.float32(reflect), np.float32(transparent)
    total = np.float32(wd + wr) + wt
    td = np.float32(wd / total)
    tr = np.float32(td + np.float32(wr / total))
    tt = np.float32(tr + np.float32(wt / total))
    m["diffuse_threshold"] = td
    m["reflect_threshold"] = tr
    m["transparent_threshold"] = tt
    m["refraction_index"] = np.float32(refraction_index)
    m["light_source"] = 1 if light else 0
    return m


class _Builder:
    def __init__(self):
        self.pos, self.nrm, self.idx, self.mat, self.obj = [], [], [], [], []
        self.nv = 0
        self.nobj = 0
        self._in_box = False

    def quad(self, corner, ea, eb, normal, m):
        """Two triangles (c, c+a, c+a+b), (c, c+a+b, c+b), flat normal; its
        own object unless part of a box."""
        if not self._in_box:
            self.nobj += 1
        self.obj += [self.nobj - 1] * 2
        c = np.asarray(corner, np.float32)
        a = np.asarray(ea, np.float32)
        b = np.asarray(eb, np.float32)
        vs = [c, c + a, c + a + b, c + b]
        for v in vs:
            self.pos.append(v)
            self.nrm.append(np.asarray(normal, np.float32))
        n = self.nv
        self.idx += [(n, n + 1, n + 2), (n, n + 2, n + 3)]
        self.mat += [m, m]
        self.nv += 4

    def box(self, lo, hi, m):
        """Axis-aligned box with outward normals, 12 triangles."""
        x0, y0, z0 = lo
        x1, y1, z1 = hi
        dx, dy, dz = x1 - x0, y1 - y0, z1 - z0
        self.nobj += 1
        self._in_box = True
        self.quad((x0, y0, z0), (0, 0, dz), (0, dy, 0), (-1, 0, 0), m)
        self.quad((x1, y0, z0), (0, dy, 0), (0, 0, dz), (1, 0, 0), m)
        self.quad((x0, y0, z0), (dx, 0, 0), (0, 0, dz), (0, -1, 0), m)
        self.quad((x0, y1, z0), (0, 0, dz), (dx, 0, 0), (0, 1, 0), m)
        self.quad((x0, y0, z0), (0, dy, 0), (dx, 0, 0), (0, 0, -1), m)
        self.quad((x0, y0, z1), (dx, 0, 0), (0, dy, 0), (0, 0, 1), m)
        self._in_box = False

    def scene(self, materials) -> Scene:
        s = Scene(
            pos=np.asarray(self.pos, np.float32).reshape(-1, 3),
            nrm=np.asarray(self.nrm, np.float32).reshape(-1, 3),
            idx=np.asarray(self.idx, np.uint32).reshape(-1, 3),
            tri_mat=np.asarray(self.mat, np.uint32),
            materials=np.stack(materials),
            tri_obj=np.asarray(self.obj, np.uint32) if len(self.obj) == len(self.mat) else None,
        )
        s.validate()
        return s


WHITE = (0.73, 0.73, 0.73)
RED = (0.65, 0.05, 0.05)
GREEN = (0.12, 0.45, 0.15)
CORNELL_HALF = 2.75


def cornell_box(light_color=(15.0, 15.0, 15.0)) -> Scene:
    """Cornell-style box seen from +x through its open side (camera on the +x
    axis looking -x, as the reference's default camera looks).  Inward wall
    normals.  Materials: 0 white, 1 red, 2 green, 3 light."""
    h = CORNELL_HALF
    b = _Builder()
    W, R, G, L = 0, 1, 2, 3
    s = 2 * h
    b.quad((-h, -h, -h), (0, 0, s), (0, s, 0), (1, 0, 0), W)      # back wall x = -h
    b.quad((-h, -h, -h), (s, 0, 0), (0, 0, s), (0, 1, 0), W)      # floor y = -h
    b.quad((-h, h, -h), (0, 0, s), (s, 0, 0), (0, -1, 0), W)      # ceiling y = +h
    b.quad((-h, -h, -h), (0, s, 0), (s, 0, 0), (0, 0, 1), G)      # wall z = -h (green)
    b.quad((-h, -h, h), (s, 0, 0), (0, s, 0), (0, 0, -1), R)      # wall z = +h (red)
    b.box((0.10, -h, 0.45), (1.75, -h + 1.65, 2.10), W)           # short box
    b.box((-1.85, -h, -2.05), (-0.20, -h + 3.30, -0.40), W)       # tall box
    lh = 0.70
    b.quad((-lh, h - 0.01, -lh), (0, 0, 2 * lh), (2 * lh, 0, 0), (0, -1, 0), L)  # light
    mats = [material(WHITE), material(RED), material(GREEN), material(light_color, light=True)]
    return b.scene(mats)


def cornell_closed(light_color=(15.0, 15.0, 15.0)) -> Scene:
    """The Cornell-style box with its open side closed by a white wall
    (x = +h, facing in): no path escapes, so paths run to the depth limit.
    Seen from inside (closed_camera).  Materials as cornell_box."""
    sc = cornell_box(light_color)
    h = CORNELL_HALF
    s = 2 * h
    b = _Builder()
    b.quad((h, -h, -h), (0, s, 0), (0, 0, s), (-1, 0, 0), 0)
    n0 = sc.pos.shape[0]
    return Scene(pos=np.concatenate([sc.pos, np.asarray(b.pos, np.float32)]),
                 nrm=np.concatenate([sc.nrm, np.asarray(b.nrm, np.float32)]),
                 idx=np.concatenate([sc.idx, np.asarray(b.idx, np.uint32) + n0]),
                 tri_mat=np.concatenate([sc.tri_mat, np.asarray(b.mat, np.uint32)]),
                 materials=sc.materials)


SPECULAR_WEIGHTS = np.array([
    # diffuse, reflect, transparent weight, refraction index per material
    (1.0, 0.0, 0.0, 1.0),   # 0 white walls
    (1.0, 0.0, 0.0, 1.0),   # 1 red
    (1.0, 0.0, 0.0, 1.0),   # 2 green
    (1.0, 0.0, 0.0, 1.0),   # 3 light
    (0.0, 1.0, 0.0, 1.0),   # 4 mirror box
    (0.0, 0.0, 1.0, 1.5),   # 5 glass ball
    (0.5, 0.3, 0.2, 1.33),  # 6 mixed ball: all three branches
], np.float32)


def cornell_specular(light_color=(15.0, 15.0, 15.0), ball_level: int = 3) -> Scene:
    """The Cornell-style box with the tall box as a mirror (reflect weight 1),
    a glass ball (transparent weight 1, refraction index 1.5, as the
    reference's cornell_box_with_ball.scene sets its ball) and a ball mixing
    all three branches (weights 0.5 / 0.3 / 0.2, index 1.33).  The balls are
    smooth-shaded icospheres (vertex normals = unit positions).  Materials
    and their weights: SPECULAR_WEIGHTS."""
    h = CORNELL_HALF
    b = _Builder()
    W, R, G, L, MIRROR, GLASS, MIXED = range(7)
    s = 2 * h
    b.quad((-h, -h, -h), (0, 0, s), (0, s, 0), (1, 0, 0), W)      # back wall x = -h
    b.quad((-h, -h, -h), (s, 0, 0), (0, 0, s), (0, 1, 0), W)      # floor y = -h
    b.quad((-h, h, -h), (0, 0, s), (s, 0, 0), (0, -1, 0), W)      # ceiling y = +h
    b.quad((-h, -h, -h), (0, s, 0), (s, 0, 0), (0, 0, 1), G)      # wall z = -h (green)
    b.quad((-h, -h, h), (s, 0, 0), (0, s, 0), (0, 0, -1), R)      # wall z = +h (red)
    b.box((-1.85, -h, -2.05), (-0.20, -h + 3.30, -0.40), MIRROR)  # tall box
    lh = 0.70
    b.quad((-lh, h - 0.01, -lh), (0, 0, 2 * lh), (2 * lh, 0, 0), (0, -1, 0), L)  # light
    v, f = icosphere(ball_level)
    for centre, radius, m in (((0.9, -h + 0.9, 1.2), 0.9, GLASS), ((0.6, -h + 0.7, -1.3), 0.7, MIXED)):
        n0 = b.nv
        for p in v:
            b.pos.append((np.asarray(centre) + radius * p).astype(np.float32))
            b.nrm.append(p.astype(np.float32))
        b.idx += [tuple(int(x) + n0 for x in t) for t in f]
        b.mat += [m] * len(f)
        b.obj += [b.nobj] * len(f)
        b.nobj += 1
        b.nv += len(v)
    mats = [material(WHITE), material(RED), material(GREEN), material(light_color, light=True)]
    mats += [material(WHITE, *SPECULAR_WEIGHTS[MIRROR][:3]),
             material((1.0, 1.0, 1.0), *SPECULAR_WEIGHTS[GLASS][:3], refraction_index=SPECULAR_WEIGHTS[GLASS][3]),
             material((0.8, 0.6, 0.3), *SPECULAR_WEIGHTS[MIXED][:3], refraction_index=SPECULAR_WEIGHTS[MIXED][3])]
    return b.scene(mats)


def cornell_indirect(light_color=(15.0, 15.0, 15.0)) -> Scene:
    """The Cornell-style box lit only indirectly: the light is a quad facing
    the ceiling on top of a white slab (the lamp) 0.8 below it, so the
    visible surfaces other than the ceiling see the light only after a
    bounce.  Materials as cornell_box."""
    h = CORNELL_HALF
    b = _Builder()
    W, R, G, L = 0, 1, 2, 3
    s = 2 * h
    b.quad((-h, -h, -h), (0, 0, s), (0, s, 0), (1, 0, 0), W)      # back wall x = -h
    b.quad((-h, -h, -h), (s, 0, 0), (0, 0, s), (0, 1, 0), W)      # floor y = -h
    b.quad((-h, h, -h), (0, 0, s), (s, 0, 0), (0, -1, 0), W)      # ceiling y = +h
    b.quad((-h, -h, -h), (0, s, 0), (s, 0, 0), (0, 0, 1), G)      # wall z = -h (green)
    b.quad((-h, -h, h), (s, 0, 0), (0, s, 0), (0, 0, -1), R)      # wall z = +h (red)
    b.box((0.10, -h, 0.45), (1.75, -h + 1.65, 2.10), W)           # short box
    b.box((-1.85, -h, -2.05), (-0.20, -h + 3.30, -0.40), W)       # tall box
    y0 = h - 0.9
    b.box((-0.8, y0, -0.8), (0.8, y0 + 0.1, 0.8), W)              # lamp slab
    lh = 0.70
    b.quad((-lh, y0 + 0.102, -lh), (0, 0, 2 * lh), (2 * lh, 0, 0), (0, 1, 0), L)  # light, facing up
    mats = [material(WHITE), material(RED), material(GREEN), material(light_color, light=True)]
    return b.scene(mats)


def cornell_groups(albedo: np.ndarray | None = None, light_color=(15.0, 15.0, 15.0), seed: int = 0) -> Scene:
    """The Cornell-style box with 64 triangle groups, each its own material
    with a learnable albedo: every wall is a 2 x 2 grid of quads (5 walls x 8
    triangles = 40) plus the two boxes (24 triangles).  Material 64 is the
    light.  albedo: (64, 3); default U(0.2, 0.8) drawn from `seed`."""
    if albedo is None:
        albedo = np.random.default_rng(seed).uniform(0.2, 0.8, (64, 3))
    albedo = np.asarray(albedo, np.float32)
    assert albedo.shape == (64, 3)
    h = CORNELL_HALF
    b = _Builder()
    g = [0]

    def grid(corner, ea, eb, normal):
        c, a, bb = (np.asarray(x, np.float32) for x in (corner, ea, eb))
        for i in range(2):
            for j in range(2):
                b.quad(c + a * (i / 2) + bb * (j / 2), a / 2, bb / 2, normal, 0)
                b.mat[-2] = g[0]
                b.mat[-1] = g[0] + 1
                g[0] += 2

    s = 2 * h
    grid((-h, -h, -h), (0, 0, s), (0, s, 0), (1, 0, 0))
    grid((-h, -h, -h), (s, 0, 0), (0, 0, s), (0, 1, 0))
    grid((-h, h, -h), (0, 0, s), (s, 0, 0), (0, -1, 0))
    grid((-h, -h, -h), (0, s, 0), (s, 0, 0), (0, 0, 1))
    grid((-h, -h, h), (s, 0, 0), (0, s, 0), (0, 0, -1))
    first_box = len(b.mat)
    b.box((0.10, -h, 0.45), (1.75, -h + 1.65, 2.10), 0)
    b.box((-1.85, -h, -2.05), (-0.20, -h + 3.30, -0.40), 0)
    for k in range(first_box, len(b.mat)):
        b.mat[k] = g[0]
        g[0] += 1
    assert g[0] == 64
    lh = 0.70
    b.quad((-lh, h - 0.01, -lh), (0, 0, 2 * lh), (2 * lh, 0, 0), (0, -1, 0), 64)
    mats = [material(albedo[k]) for k in range(64)] + [material(light_color, light=True)]
    return b.scene(mats)


def closed_room(floor_mat: np.ndarray, wall_mat: np.ndarray, half: float = 1.0) -> Scene:
    """Closed cube [-half, half]^3 seen from inside (inward normals): the
    floor y = -half uses material 0, the other five faces material 1.  Used
    by the estimator's closed-form checks."""
    h = half
    s = 2 * h
    b = _Builder()
    b.quad((-h, -h, -h), (s, 0, 0), (0, 0, s), (0, 1, 0), 0)      # floor
    b.quad((-h, h, -h), (0, 0, s), (s, 0, 0), (0, -1, 0), 1)      # ceiling
    b.quad((-h, -h, -h), (0, 0, s), (0, s, 0), (1, 0, 0), 1)      # x = -h
    b.quad((h, -h, -h), (0, s, 0), (0, 0, s), (-1, 0, 0), 1)      # x = +h
    b.quad((-h, -h, -h), (0, s, 0), (s, 0, 0), (0, 0, 1), 1)      # z = -h
    b.quad((-h, -h, h), (s, 0, 0), (0, s, 0), (0, 0, -1), 1)      # z = +h
    return b.scene([floor_mat, wall_mat])


def ground_plane(albedo, half: float = 100.0) -> Scene:
    """One upward-facing square y = 0 of side 2 * half (material 0): open
    sky above it.  Used by the sun's closed-form checks."""
    b = _Builder()
    s = 2 * half
    b.quad((-half, 0.0, -half), (0, 0, s), (s, 0, 0), (0, 1, 0), 0)
    return b.scene([material(albedo)])


def cornell_camera_position() -> tuple[float, float, float]:
    """Eye on the +x axis far enough for a 40-degree field of view to frame the
    box's open face."""
    return (CORNELL_HALF + CORNELL_HALF / np.tan(np.radians(20.0)) + 0.05, 0.0, 0.0)


def _hash01(ix: np.ndarray, iz: np.ndarray, octave: int, seed: int) -> np.ndarray:
    """Lattice value in [0, 1) from a splitmix64-style integer hash."""
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ (iz.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F))
        h ^= np.uint64((octave * 0x165667B19E3779F9 + seed * 0x27D4EB2F165667C5) & 0xFFFFFFFFFFFFFFFF)
        h ^= h >> np.uint64(30)
        h *= np.uint64(0xBF58476D1CE4E5B9)
        h ^= h >> np.uint64(27)
        h *= np.uint64(0x94D049BB133111EB)
        h ^= h >> np.uint64(31)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def fbm_heightfield(n: int = 2048, octaves: int = 6, seed: int = 7, base_cells: float = 8.0) -> np.ndarray:
    """n x n heights in [-1, 1]-ish: value noise (smoothstep-interpolated lattice
    hashes), lacunarity 2, gain 0.5."""
    u = np.arange(n, dtype=np.float64) / (n - 1)
    h = np.zeros((n, n), np.float64)
    amp, freq, total = 1.0, base_cells, 0.0
    for o in range(octaves):
        x = u * freq
        i0 = np.floor(x).astype(np.int64)
        f = x - i0
        s = f * f * (3 - 2 * f)
        X0, Z0 = np.meshgrid(i0, i0, indexing="xy")
        SX, SZ = np.meshgrid(s, s, indexing="xy")
        v00 = _hash01(X0, Z0, o, seed)
        v10 = _hash01(X0 + 1, Z0, o, seed)
        v01 = _hash01(X0, Z0 + 1, o, seed)
        v11 = _hash01(X0 + 1, Z0 + 1, o, seed)
        val = (v00 * (1 - SX) + v10 * SX) * (1 - SZ) + (v01 * (1 - SX) + v11 * SX) * SZ
        h += amp * (2 * val - 1)
        total += amp
        amp *= 0.5
        freq *= 2.0
    return h / total


def icosphere(level: int = 5):
    """Unit icosphere: (vertices, triangles) after `level` midpoint subdivisions
    (20 * 4^level triangles)."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts, np.float64), np.array(faces, np.int64)


def terrain_scene(n: int = 2048, spheres: int = 64, sphere_level: int = 5, extent: float = 100.0,
                  height: float = 12.0, seed: int = 7) -> Scene:
    """Procedural landscape: an n x n fBm heightfield (2 (n-1)^2 triangles,
    6 octaves) over [-extent/2, extent/2]^2 plus `spheres` icospheres at
    seed-drawn positions resting on it.  Lambertian albedos U(0.1, 0.9): one
    for the terrain, one per sphere.  Lighting comes from the sky and sun
    settings (RT_RENDER_SKY / RT_RENDER_SUN)."""
    rng = np.random.default_rng(seed)
    hf = fbm_heightfield(n, seed=seed) * height
    xs = (np.arange(n) / (n - 1) - 0.5) * extent
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    pos = np.stack([X, hf, Z], axis=-1).reshape(-1, 3)
    # normals from central differences of the height field
    dx = extent / (n - 1)
    gx = np.gradient(hf, dx, axis=1)
    gz = np.gradient(hf, dx, axis=0)
    nrm = np.stack([-gx, np.ones_like(gx), -gz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    nrm = nrm.reshape(-1, 3)
    i = np.arange(n - 1)
    I, J = np.meshgrid(i, i, indexing="xy")
    v00 = (J * n + I).reshape(-1)
    v10, v01, v11 = v00 + 1, v00 + n, v00 + n + 1
    tri = np.concatenate([np.stack([v00, v01, v10], 1), np.stack([v10, v01, v11], 1)]).astype(np.uint32)
    tri_mat = np.zeros(tri.shape[0], np.uint32)
    mats = [material(rng.uniform(0.1, 0.9, 3))]
    sv, sf = icosphere(sphere_level)
    all_pos, all_nrm, all_tri, all_mat = [pos], [nrm], [tri], [tri_mat]
    base = pos.shape[0]
    for k in range(spheres):
        r = rng.uniform(1.0, 3.0)
        cx, cz = rng.uniform(-0.4, 0.4, 2) * extent
        gi = int(round((cx / extent + 0.5) * (n - 1)))
        gj = int(round((cz / extent + 0.5) * (n - 1)))
        cy = hf[gj, gi] + 0.8 * r
        all_pos.append(sv * r + [cx, cy, cz])
        all_nrm.append(sv)
        all_tri.append((sf + base).astype(np.uint32))
        all_mat.append(np.full(sf.shape[0], k + 1, np.uint32))
        mats.append(material(rng.uniform(0.1, 0.9, 3)))
        base += sv.shape[0]
    s = Scene(pos=np.concatenate(all_pos).astype(np.float32), nrm=np.concatenate(all_nrm).astype(np.float32),
              idx=np.ascontiguousarray(np.concatenate(all_tri)), tri_mat=np.concatenate(all_mat),
              materials=np.stack(mats))
    s.validate()
    return s


def terrain_cameras(num_views: int = 16, width: int = 1920, height: int = 1080, radius: float = 42.0,
                    elevation: float = 22.0, fov_deg: float = 45.0):
    """Cameras on a circle around the terrain looking at its centre (in the
    reference's camera terms: angle0 = theta / pi, angle1 maps phi)."""
    from . import host_lib as H
    cams = []
    for v in range(num_views):
        a = 2 * np.pi * v / num_views
        p = np.array([radius * np.cos(a), elevation, radius * np.sin(a)])
        d = -p / np.linalg.norm(p)
        theta = np.arctan2(d[2], d[0])
        phi = np.arccos(np.clip(d[1], -1, 1))
        cams.append(H.camera(tuple(p), float(theta / np.pi), float((np.pi / 2 - phi) / (np.pi / 2)),
                             float(np.radians(fov_deg)), width, height))
    return cams


def _unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def triangle_soup(T: int = 1 << 20, seed: int = 1234, mu: float = -5.0, sigma: float = 0.5) -> Scene:
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, (T, 1, 3))
    length = rng.lognormal(mu, sigma, (T, 3, 1))
    dirs = _unit_vectors(rng, 3 * T).reshape(T, 3, 3)
    verts = (c + length * dirs).astype(np.float32)          # (T, 3, 3)
    e0 = verts[:, 1] - verts[:, 0]
    e1 = verts[:, 2] - verts[:, 0]
    n = np.cross(e0, e1)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    pos = verts.reshape(-1, 3)
    nrm = np.repeat(n.astype(np.float32), 3, axis=0)
    idx = np.arange(3 * T, dtype=np.uint32).reshape(T, 3)
    s = Scene(pos=np.ascontiguousarray(pos), nrm=np.ascontiguousarray(nrm), idx=idx,
              tri_mat=np.zeros(T, np.uint32), materials=np.stack([material(WHITE)]))
    s.validate()
    return s


def pinhole_rays(eye, target, fov_deg: float, width: int, height: int) -> np.ndarray:
    """One ray per pixel centre of a pinhole camera (row-major, y down)."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    up = np.array([0.0, 1.0, 0.0])
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    t = np.tan(np.radians(fov_deg) / 2)
    aspect = width / height
    px = (np.arange(width) + 0.5) / width * 2 - 1
    py = 1 - (np.arange(height) + 0.5) / height * 2
    X, Y = np.meshgrid(px * t * aspect, py * t)
    d = X[..., None] * s + Y[..., None] * u + f
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    r = np.zeros(width * height, RAY_DTYPE)
    r["o"] = eye.astype(np.float32)
    r["d"] = d.reshape(-1, 3).astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = np.inf
    return r


def incoherent_rays(n: int, seed: int = 4321) -> np.ndarray:
    """Origins uniform in the unit cube, directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY_DTYPE)
    r["o"] = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    r["d"] = _unit_vectors(rng, n).astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = np.inf
    return r


def microbench_rays(width: int = 4096, height: int = 2048, n_incoherent: int | None = None, seed: int = 4321):
    """The ray-cast microbenchmark's two halves: coherent pinhole rays from
    (0.5, 0.5, -1.5) at the cube centre, FOV 40 degrees, and as many
    incoherent rays."""
    coh = pinhole_rays((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), 40.0, width, height)
    inc = incoherent_rays(width * height if n_incoherent is None else n_incoherent, seed)
    return coh, inc
