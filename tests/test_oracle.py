"""Pins the CPU oracle independently of its own formulas: published Philox
known-answer vectors, closed forms of Moller-Trumbore and of the barycentric
coordinates, libm for the sampling angle, distribution moments, brute force
for the BVH, and closed-form means of the reference estimator."""
import math

import numpy as np
import pytest

import oracle as O
from master_thesis_amd import scenes, host_lib as H
from master_thesis_amd._lib import RAY_DTYPE, RT_MISS
from master_thesis_amd.tracer import make_settings


# ---------------------------------------------------------------- Philox
# Known-answer vectors of Philox-4x32-10 published with Random123
# (Salmon et al., SC'11): counter, key -> output.
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(int(x) for x in O.philox(ctr, key)) == out


# ---------------------------------------------------------------- sampling
def test_sincos_2pi_against_libm():
    rng = np.random.default_rng(0)
    ks = np.concatenate([rng.integers(0, 1 << 24, 20000), np.arange(0, 1 << 24, 1 << 21),
                         np.arange(0, 1 << 24, 1 << 21) - 1, [0, 1, (1 << 24) - 1]])
    ks = ks[(ks >= 0) & (ks < (1 << 24))]
    worst = 0.0
    for k in ks:
        x = np.float32(k / float(1 << 24))
        s, c = O.sincos_2pi(float(x))
        ang = 2 * math.pi * float(x)
        worst = max(worst, abs(s - math.sin(ang)), abs(c - math.cos(ang)))
        assert abs(s * s + c * c - 1.0) < 5e-7
    # |sin|, |cos| <= 1: 2 ulp of 1.0 in float is 2.4e-7
    assert worst < 2.4e-7, worst


def test_sincos_2pi_octant_symmetry():
    # exact values at multiples of 1/8 turn
    s, c = O.sincos_2pi(0.0)
    assert (s, c) == (0.0, 1.0)
    s, c = O.sincos_2pi(0.25)
    assert (s, c) == (1.0, 0.0) or (abs(c) < 1e-7 and s == 1.0)
    s, c = O.sincos_2pi(0.5)
    assert abs(s) < 1e-7 and c == -1.0


def test_hemisphere_distribution():
    """Uniform hemisphere about n: n.d ~ U[0, 1] (Archimedes), |d| = 1."""
    import ctypes as C
    lib = O.load()
    n = np.array([0.3, -0.5, 0.81], np.float32)
    n /= np.linalg.norm(n)
    rng = np.random.default_rng(1)
    out = np.zeros(3, np.float32)
    cosines = []
    for _ in range(20000):
        xi = (rng.integers(0, 1 << 24, 2) / float(1 << 24)).astype(np.float32)
        lib.or_hemisphere(float(xi[0]), float(xi[1]), O.fptr(n), O.fptr(out))
        assert abs(np.linalg.norm(out) - 1.0) < 1e-6
        cosines.append(float(out @ n))
    c = np.array(cosines)
    assert c.min() >= -1e-7
    assert abs(c.mean() - 0.5) < 4 * math.sqrt(1 / 12 / len(c))
    hist, _ = np.histogram(c, bins=10, range=(0, 1))
    assert np.all(np.abs(hist - len(c) / 10) < 5 * math.sqrt(len(c) / 10))


# ---------------------------------------------------------------- Moller-Trumbore
PANEL = (np.array([0, -1, -1], np.float32), np.array([0, 1, -1], np.float32), np.array([0, 1, 1], np.float32))


def test_mt_closed_form_hit():
    hit, (t, u, v) = O.intersect((2, 0.5, -0.5), (-1, 0, 0), *PANEL)
    assert hit and (t, u, v) == (2.0, 0.5, 0.25)


def test_mt_two_sided():
    hit, (t, u, v) = O.intersect((-2, 0.5, -0.5), (1, 0, 0), *PANEL)
    assert hit and (t, u, v) == (2.0, 0.5, 0.25)


def test_mt_parallel_and_behind_and_on_plane():
    assert not O.intersect((2, 0.5, -0.5), (0, 1, 0), *PANEL)[0]       # parallel: |a| < 1e-7
    assert not O.intersect((2, 0.5, -0.5), (1, 0, 0), *PANEL)[0]       # plane behind: t < 0
    assert not O.intersect((0, 0.5, -0.5), (-1, 0, 0), *PANEL)[0]      # origin on the plane: t = 0


@pytest.mark.parametrize("y,z", [(0.25, 0.25), (0.5, -1.0), (1.0, 0.0), (1.0, 1.0), (-1.0, -1.0)])
def test_mt_edges_inclusive(y, z):
    hit, (t, u, v) = O.intersect((2, y, z), (-1, 0, 0), *PANEL)
    assert hit and t == 2.0
    assert u == (y - z) / 2 and v == (z + 1) / 2


@pytest.mark.parametrize("y,z", [(1.0 + 2 ** -20, 0.0), (0.25, 0.25 + 2 ** -20), (0.5, -1.0 - 2 ** -20)])
def test_mt_just_outside(y, z):
    assert not O.intersect((2, y, z), (-1, 0, 0), *PANEL)[0]


def test_barycentric_closed_form():
    b = O.barycentric(*PANEL, (0, 0.5, -0.5))
    assert tuple(b) == (0.25, 0.5, 0.25)
    rng = np.random.default_rng(3)
    for _ in range(200):
        v = rng.normal(size=(3, 3)).astype(np.float32)
        w = rng.dirichlet([1, 1, 1])
        p = (w[0] * v[0] + w[1] * v[1] + w[2] * v[2]).astype(np.float32)
        b = O.barycentric(v[0], v[1], v[2], p)
        assert abs(float(b.sum()) - 1.0) < 1e-5
        assert np.allclose(b, w, atol=2e-3)


# ---------------------------------------------------------------- reflect / refract
def _unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def test_reflect_closed_form():
    """Mirror about the normal: the normal component flips, the tangential
    one stays; reflecting twice gives the input back (within rounding)."""
    I = _unit((1, -1, 0))
    r = O.reflect(I, (0, 1, 0))
    assert np.allclose(r, _unit((1, 1, 0)), atol=1e-7)
    rng = np.random.default_rng(5)
    for _ in range(500):
        I, N = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        r = O.reflect(I, N)
        assert abs(float(np.dot(r, N)) + float(np.dot(I, N))) < 1e-6
        assert np.allclose(r - np.dot(r, N) * N, I - np.dot(I, N) * N, atol=1e-6)
        assert np.allclose(O.reflect(r, N), I, atol=1e-6)


def test_custom_refract_snell_and_total_internal_reflection():
    """customRefract (math/vector.h:332-345): entering (n.i < 0) bends with
    sin t = sin i / eta; leaving (n.i > 0) the normal is flipped and
    sin t = eta sin i, or, past the critical angle sin i >= 1 / eta, the ray
    is reflected.  The refracted ray is unit length, stays in the plane of
    incidence and crosses the surface; a swapped normal flip or a missing
    1 / eta fails these."""
    rng = np.random.default_rng(6)
    n_refract = n_tir = 0
    for _ in range(3000):
        N = _unit(rng.normal(size=3))
        I = _unit(rng.normal(size=3))
        eta = np.float32(rng.uniform(1.05, 2.4))
        c = float(np.dot(N, I))
        sin_i = math.sqrt(max(0.0, 1 - c * c))
        out = O.custom_refract(I, N, eta)
        ratio = (1 / float(eta)) if c <= 0 else float(eta)
        if sin_i * ratio < 1 - 1e-5:
            n_refract += 1
            assert abs(np.linalg.norm(out) - 1) < 2e-6
            side = -1.0 if c <= 0 else 1.0          # the refracted ray keeps going through
            assert np.sign(np.dot(out, N)) == side
            sin_t = np.linalg.norm(np.cross(out, N))
            assert abs(sin_t - sin_i * ratio) < 2e-5
            assert abs(np.dot(np.cross(I, N), out)) < 2e-5  # plane of incidence
        elif sin_i * ratio > 1 + 1e-5:
            n_tir += 1
            assert c > 0
            assert np.array_equal(out, O.reflect(I, N))
    assert n_refract > 1500 and n_tir > 200
    # normal incidence passes straight through either way
    assert np.allclose(O.custom_refract((0, -1, 0), (0, 1, 0), np.float32(1.5)), (0, -1, 0), atol=1e-7)
    assert np.allclose(O.custom_refract((0, 1, 0), (0, 1, 0), np.float32(1.5)), (0, 1, 0), atol=1e-7)


def test_sin_acos_against_libm():
    """customRefract's sin(acos(x)) as sqrt((1 - x)(1 + x)): within 2^-50 of
    glibc's sin(acos(x)) for float x over [-1, 1], dense near both ends where
    acos is ill-conditioned; NaN outside like acos."""
    rng = np.random.default_rng(7)
    one = np.float32(1.0)
    xs = np.concatenate([rng.uniform(-1, 1, 50000).astype(np.float32),
                         one - np.arange(4000, dtype=np.float32) * np.float32(2 ** -24),
                         -one + np.arange(4000, dtype=np.float32) * np.float32(2 ** -24),
                         np.float32([0.0, -0.0, 0.5, -0.5, 2 ** -30])])
    worst = max(abs(O.sin_acos(float(x)) - math.sin(math.acos(float(x)))) for x in xs)
    assert worst <= 2 ** -50, worst
    assert O.sin_acos(1.0) == 0.0
    assert math.isnan(O.sin_acos(1.0000001)) and math.isnan(O.sin_acos(-1.0000001))


# ---------------------------------------------------------------- closest hit
def _rays(o, d):
    r = np.zeros(len(o), RAY_DTYPE)
    r["o"] = o
    r["d"] = d
    r["tmax"] = np.inf
    return r


def test_brute_force_picks_nearest_of_stacked_panels():
    """Three parallel quads at x = 0, 1, 2; rays from x = 5 looking -x hit
    x = 2 first; the reference criterion is the squared distance."""
    b = scenes._Builder()
    for x in (0.0, 2.0, 1.0):
        b.quad((x, -1, -1), (0, 2, 0), (0, 0, 2), (1, 0, 0), 0)
    sc = b.scene([scenes.material(scenes.WHITE)])
    rng = np.random.default_rng(5)
    o = np.column_stack([np.full(100, 5.0), rng.uniform(-0.9, 0.9, (100, 2))]).astype(np.float32)
    d = np.tile(np.array([-1, 0, 0], np.float32), (100, 1))
    h = O.trace_brute(sc, _rays(o, d))
    assert np.all((h["tri"] == 2) | (h["tri"] == 3))
    assert np.all(h["t"] == 3.0)


def test_tie_goes_to_lowest_index():
    """Two identical triangles: exact squared-distance tie -> lower index."""
    b = scenes._Builder()
    b.quad((0, -1, -1), (0, 2, 0), (0, 0, 2), (1, 0, 0), 0)
    b.quad((0, -1, -1), (0, 2, 0), (0, 0, 2), (1, 0, 0), 0)
    sc = b.scene([scenes.material(scenes.WHITE)])
    h = O.trace_brute(sc, _rays(np.array([[3, 0.5, 0.25]], np.float32), np.array([[-1, 0, 0]], np.float32)))
    assert h["tri"][0] == 0


@pytest.mark.parametrize("T,seed", [(1, 0), (7, 1), (500, 2), (20000, 3)])
def test_cpu_bvh_equals_brute_force(T, seed):
    sc = scenes.triangle_soup(T, seed=seed, mu=-2.5 if T < 100 else -4.0)
    coh, inc = scenes.microbench_rays(48, 24, 1500, seed=seed)
    rays = np.concatenate([coh, inc])
    hb = O.trace_brute(sc, rays)
    hv, cnt = O.CpuBVH(sc).trace(rays)
    assert np.array_equal(hb.view(np.uint32), hv.view(np.uint32))
    assert cnt["rays"] == rays.shape[0]


def test_cpu_bvh_degenerate_coincident_centroids():
    """All triangles share one centroid: the build splits by index."""
    rng = np.random.default_rng(9)
    T = 300
    d = rng.normal(size=(T, 3, 3)).astype(np.float32) * 0.1
    d -= d.mean(axis=1, keepdims=True)
    pos = (d + 0.5).reshape(-1, 3).astype(np.float32)
    sc = scenes.Scene(pos=pos, nrm=np.zeros_like(pos), idx=np.arange(3 * T, dtype=np.uint32).reshape(T, 3),
                      tri_mat=np.zeros(T, np.uint32), materials=np.stack([scenes.material(scenes.WHITE)]))
    rays = scenes.incoherent_rays(3000, seed=4)
    hb = O.trace_brute(sc, rays)
    hv, _ = O.CpuBVH(sc, leaf_size=2).trace(rays)
    assert np.array_equal(hb.view(np.uint32), hv.view(np.uint32))


# ---------------------------------------------------------------- estimator
def _room_camera(w=32, h=32):
    # inside the room at the centre, looking straight down (angle1 = -1 -> phi clamped to pi - 0.01)
    return H.camera((0.0, 0.0, 0.0), 0.0, -1.0, math.pi / 4, w, h)


def test_estimator_all_light_room_is_pi_E():
    E = 1.5
    light = scenes.material((E, E, E), light=True)
    sc = scenes.closed_room(light, light)
    cam = _room_camera()
    N = 16
    rgb, _ = O.render(sc, cam, make_settings(spp=N, max_depth=5, seed=11))
    P = cam.width * cam.height
    sd = 2 * math.pi * E / math.sqrt(12 * N * P)
    assert abs(rgb.mean() - math.pi * E) < 4 * sd


def test_estimator_lit_floor_is_half_pi_rho_E():
    E, rho = 2.0, 0.6
    sc = scenes.closed_room(scenes.material((rho, rho, rho)), scenes.material((E, E, E), light=True))
    cam = _room_camera()
    N = 32
    rgb, _ = O.render(sc, cam, make_settings(spp=N, max_depth=5, seed=12))
    P = cam.width * cam.height
    # per-sample value 2*pi*rho*E*c1*c2, c1, c2 ~ U[0,1]: variance (2 pi rho E)^2 (1/9 - 1/16)
    sd = 2 * math.pi * rho * E * math.sqrt(1 / 9 - 1 / 16) / math.sqrt(N * P)
    assert abs(rgb.mean() - math.pi * rho * E / 2) < 4 * sd


def test_estimator_no_lights_is_black():
    w = scenes.material(scenes.WHITE)
    sc = scenes.closed_room(w, w)
    rgb, rgb8 = O.render(sc, _room_camera(16, 16), make_settings(spp=4, max_depth=5, seed=1))
    assert not rgb.any() and not rgb8.any()


def test_estimator_depth_one_only_direct_light_hits():
    """max_depth 1: only primary rays that land on a light contribute, with
    value 2*pi/N * sum(E * cos)."""
    E = 1.0
    light = scenes.material((E, E, E), light=True)
    sc = scenes.closed_room(scenes.material(scenes.WHITE), light)
    cam = _room_camera(16, 16)
    rgb, _ = O.render(sc, cam, make_settings(spp=8, max_depth=1, seed=3))
    assert not rgb.any()   # every primary ray hits the (non-light) floor


def test_rgb8_is_clamped_truncation():
    E = 1.5
    light = scenes.material((E, E, E), light=True)
    sc = scenes.closed_room(light, light)
    rgb, rgb8 = O.render(sc, _room_camera(8, 8), make_settings(spp=2, max_depth=3, seed=2))
    expect = (np.clip(rgb, 0, 1) * np.float32(255.0)).astype(np.int32).astype(np.uint8)
    assert np.array_equal(rgb8, expect)


def test_bidirectional_equals_path_tracer_when_paths_end_on_lights():
    """Floor under emissive walls: every vision path ends on a light at its
    second vertex, so the bidirectional tracer adds exactly the path
    tracer's contribution (same Philox counters on the vision side): the
    images are identical, and their mean is pi * rho * E / 2."""
    E, rho = 2.0, 0.6
    sc = scenes.closed_room(scenes.material((rho, rho, rho)), scenes.material((E, E, E), light=True))
    cam = _room_camera(24, 24)
    s = make_settings(spp=16, max_depth=5, seed=12)
    pt, _ = O.render(sc, cam, s)
    bd = O.render_bdpt(sc, cam, s, vision_jumps=5, light_jumps=5, max_depth=5)
    np.testing.assert_array_equal(bd, pt)



def test_faithful_mode_matches_oracle_estimator():
    """The faithful-mode CPU baseline (mt19937 stream, the reference's BVH
    queries) estimates the oracle's image (Philox stream, binned-SAH BVH):
    block means within the spread across oracle seeds; on an all-light room
    it gives pi * E like the oracle; a first-pixels run touches only those
    pixels."""
    sc = scenes.cornell_box()
    cam = H.camera(scenes.cornell_camera_position(), 1.0, 0.0, np.float32(math.pi / 4), 32, 32)
    img = O.render_faithful(sc, cam, 256, 5, threads=4, seed=5)[0]
    runs = np.stack([O.render(sc, cam, make_settings(spp=256, max_depth=5, seed=s))[0] for s in range(1, 13)])
    b = img.reshape(4, 8, 4, 8, 3).mean(axis=(1, 3))
    ob = runs.reshape(12, 4, 8, 4, 8, 3).mean(axis=(2, 4))
    z = (b - ob.mean(axis=0)) / np.maximum(ob.std(axis=0, ddof=1) * math.sqrt(1 + 1 / 12), 1e-6)
    assert np.all(np.abs(z) < 7), np.abs(z).max()
    E = 1.5
    light = scenes.material((E, E, E), light=True)
    room = scenes.closed_room(light, light)
    rc = H.camera((0.0, 0.0, 0.0), 0.0, -1.0, math.pi / 4, 32, 32)
    im = O.render_faithful(room, rc, 64, 5, threads=4, seed=2)[0]
    sd = 2 * math.pi * E / math.sqrt(12 * 64 * 32 * 32)
    assert abs(im.mean() - math.pi * E) < 5 * sd
    part = O.render_faithful(sc, cam, 4, 5, threads=2, seed=3, pixel_limit=600)[0].reshape(-1, 3)
    assert np.all(part[600:] == 0) and part[:600].max() > 0


def _quantize_ref(lo, hi, valid):
    """Outward 8-bit quantisation of child boxes (lo, hi: (4, 3)) against
    their union, as rt_bvh4q_node stores it (restated here in numpy)."""
    from master_thesis_amd._lib import QNODE_DTYPE
    q = np.zeros((), QNODE_DTYPE)
    for k, ax in enumerate("xyz"):
        o = np.float32(lo[valid, k].min())
        ext = float(hi[valid, k].max()) - float(o)
        e = math.frexp(ext / 255.0)[1] if ext > 0 else -100
        if math.ldexp(1.0, e - 1) >= ext / 255.0:
            e -= 1
        q["origin"][k] = o
        q["exp"][k] = e
        for c in range(4):
            a = math.floor(math.ldexp(float(lo[c, k]) - float(o), -e)) if valid[c] else 255
            b = math.ceil(math.ldexp(float(hi[c, k]) - float(o), -e)) if valid[c] else 0
            q["lo_" + ax][c] = min(max(a, 0), 255)
            q["hi_" + ax][c] = min(max(b, 0), 255)
    return q


def test_quantized_node_check_catches_inward_rounding():
    """or_check_quantized accepts outward-rounded 8-bit child boxes and
    rejects a node whose low bound was rounded to nearest (it can cut a
    sliver off the float box), a changed child, or an empty child's code."""
    from master_thesis_amd._lib import NODE_DTYPE, QNODE_DTYPE
    rng = np.random.default_rng(9)
    nodes = np.zeros(200, NODE_DTYPE)
    qn = np.zeros(200, QNODE_DTYPE)
    for i in range(200):
        valid = np.array([True, True, i % 3 != 0, i % 5 != 0])
        c = rng.uniform(-10, 10, (4, 3)).astype(np.float32)
        ext = rng.lognormal(-2, 2, (4, 3)).astype(np.float32)
        lo, hi = (c - ext).astype(np.float32), (c + ext).astype(np.float32)
        for k, ax in enumerate("xyz"):
            nodes[i]["lo_" + ax] = lo[:, k]
            nodes[i]["hi_" + ax] = hi[:, k]
        nodes[i]["child"] = np.where(valid, np.arange(4) + 4 * i + 1, 0xFFFFFFFF)
        nodes[i]["meta"][0] = i // 4
        qn[i] = _quantize_ref(lo, hi, valid)
        qn[i]["child"] = nodes[i]["child"]
        qn[i]["parent"] = i // 4
    assert O.check_quantized(qn, nodes) == 0
    bad = qn.copy()
    for i in range(200):  # round the first child's low x bound to nearest
        o, e = float(bad[i]["origin"][0]), int(bad[i]["exp"][0])
        bad[i]["lo_x"][0] = min(255, round(math.ldexp(float(nodes[i]["lo_x"][0]) - o, -e)))
    assert O.check_quantized(bad, nodes) == 4
    bad = qn.copy()
    bad[1]["child"][0] ^= 1
    assert O.check_quantized(bad, nodes) == 2
    bad = qn.copy()
    bad[0]["lo_y"][2] = 0
    assert O.check_quantized(bad, nodes) == 3
