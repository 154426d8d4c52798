"""The oracle's irradiance-probe restatement (rt_ddgi_*: the Vulkan renderer's
Majercik2019 pipeline in the path tracer's units), pinned by closed forms:
spherical Fibonacci directions, octahedral folding, the cosine-weighted
update, Chebyshev visibility, trilinear blending, the sun's closed form at
probe-ray hits and in the render."""
import math

import numpy as np
import pytest

import oracle as O
from master_thesis_amd import _lib as L
from master_thesis_amd import host_lib as H
from master_thesis_amd import scenes
from master_thesis_amd.tracer import make_settings


def ddgi_cfg(count=(2, 2, 2), start=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0), rays=64, side=8, max_dist=10.0,
             sharpness=0.8, bias=0.0, energy=1.0, hysteresis=0.0, shadow_count=1, flags=0):
    c = L.rt_ddgi_config()
    for k in range(3):
        c.probe_count[k] = count[k]
        c.probe_start[k] = start[k]
        c.probe_spacing[k] = spacing[k]
    c.rays_per_probe = rays
    c.side = side
    c.max_ray_distance = max_dist
    c.depth_sharpness = sharpness
    c.normal_bias = bias
    c.energy_preservation = energy
    c.hysteresis = hysteresis
    c.shadow_count = shadow_count
    c.flags = flags
    return c


def fib_dirs(n):
    return np.array([O.ddgi_fib(i, n) for i in range(n)])


def surfels_from(dirs, rad, dist=None, hit=None, probes=1):
    n = dirs.shape[0]
    s = np.zeros((probes * n, 8), np.float32)
    for p in range(probes):
        s[p * n:(p + 1) * n, 0:3] = dirs
        s[p * n:(p + 1) * n, 3:6] = rad
        s[p * n:(p + 1) * n, 6] = 1.0 if dist is None else dist
        s[p * n:(p + 1) * n, 7] = 0.0 if hit is None else hit
    return s


def test_spherical_fibonacci_closed_form():
    """cos(theta_i) = 1 - (2 i + 1) / n and phi_i = 2 pi frac(i sqrt5 / 2 -
    1/2), against float64 libm (the float product i sqrt5 / 2 carries an
    angle error of up to 2 pi ulp(i sqrt5 / 2)); unit length; the set's
    centroid near 0."""
    n = 256
    d = fib_dirs(n)
    for i in range(n):
        ct = 1.0 - (2 * i + 1) / n
        ph = 2 * math.pi * ((i * math.sqrt(5) * 0.5 - 0.5) % 1.0)
        st = math.sqrt(1 - ct * ct)
        tol = 2e-6 + 2 * math.pi * float(np.spacing(np.float32(i * math.sqrt(5) * 0.5 + 1.0)))
        np.testing.assert_allclose(d[i], [math.cos(ph) * st, math.sin(ph) * st, ct], atol=tol)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.abs(d.mean(axis=0)).max() < 2.0 / n


def test_octahedral_round_trip_and_folds():
    rng = np.random.default_rng(1)
    v = rng.normal(size=(500, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for x in v:
        o = O.ddgi_oct_encode(x)
        assert abs(o[0]) <= 1 and abs(o[1]) <= 1
        # upper hemisphere inside the diamond, lower outside
        assert (abs(o[0]) + abs(o[1]) <= 1 + 1e-6) == (x[2] >= -1e-7) or abs(x[2]) < 1e-6
        np.testing.assert_allclose(O.ddgi_oct_decode(o), x, atol=2e-6)
    np.testing.assert_allclose(O.ddgi_oct_encode([0, 0, 1]), [0, 0], atol=0)
    np.testing.assert_allclose(O.ddgi_oct_encode([1, 0, 0]), [1, 0], atol=0)
    np.testing.assert_allclose(O.ddgi_oct_encode([0, 0, -1]), [1, 1], atol=0)
    # the map folds at its edges: (+-1, y) and (+-1, -y) are one direction, as
    # are (x, +-1) and (-x, +-1); all four corners are -z (the tile border's rule)
    for y in np.linspace(-1, 1, 9):
        for s in (-1.0, 1.0):
            np.testing.assert_allclose(O.ddgi_oct_decode([s, y]), O.ddgi_oct_decode([s, -y]), atol=1e-7)
            np.testing.assert_allclose(O.ddgi_oct_decode([y, s]), O.ddgi_oct_decode([-y, s]), atol=1e-7)
    for cx in (-1, 1):
        for cy in (-1, 1):
            np.testing.assert_allclose(O.ddgi_oct_decode([cx, cy]), [0, 0, -1], atol=1e-7)


def test_update_constant_radiance_and_distance():
    """Every texel of every probe: irradiance = the constant radiance, depth
    moments (r, r^2) whatever the weights; misses sit at max distance."""
    cfg = ddgi_cfg(count=(2, 1, 2), rays=128, max_dist=7.5)
    d = fib_dirs(128)
    s = surfels_from(d, [0.3, 1.5, 4.0], dist=2.25, hit=1.0, probes=4)
    irr, dep = O.ddgi_update(cfg, s)
    assert irr.shape == (2 * 10, 2 * 10, 4)
    np.testing.assert_allclose(irr[..., :3], np.broadcast_to([0.3, 1.5, 4.0], irr[..., :3].shape), rtol=2e-6)
    np.testing.assert_allclose(dep[..., 0], 2.25, rtol=1e-6)
    np.testing.assert_allclose(dep[..., 1], 2.25 ** 2, rtol=1e-6)
    s[:, 7] = 0.0  # all misses
    _, dep = O.ddgi_update(cfg, s)
    np.testing.assert_allclose(dep[..., 0], 7.5, rtol=1e-6)
    s[:, 7] = 1.0
    s[:, 6] = 100.0  # hits beyond the cap count at the cap
    _, dep = O.ddgi_update(cfg, s)
    np.testing.assert_allclose(dep[..., 1], 7.5 ** 2, rtol=1e-6)


def test_update_is_the_cosine_weighted_mean():
    """Radiance a . d (linear in the ray direction): the cosine-weighted mean
    over the hemisphere about the texel direction n is (2/3) a . n (uniform
    weights would give 1/2, cos^2 weights 3/4)."""
    R, side = 1024, 8
    cfg = ddgi_cfg(count=(1, 1, 1), rays=R, side=side)
    d = fib_dirs(R)
    a = np.array([0.8, -0.5, 0.3])
    rad = np.stack([d @ a] * 3, axis=1)
    irr, _ = O.ddgi_update(cfg, surfels_from(d, rad))
    for ix in range(1, side + 1):
        for iy in range(1, side + 1):
            n = O.ddgi_oct_decode([(ix - 1 + 0.5) * 2 / side - 1, (iy - 1 + 0.5) * 2 / side - 1])
            assert abs(irr[iy, ix, 0] - (2.0 / 3.0) * float(a @ n)) < 0.01


def test_update_border_holds_the_folded_neighbours():
    """A border texel holds the interior texel across the fold (same
    direction on the sphere as the texel just outside the edge); random
    surfels."""
    side = 6
    cfg = ddgi_cfg(count=(1, 1, 1), rays=64, side=side)
    rng = np.random.default_rng(3)
    s = surfels_from(fib_dirs(64), rng.uniform(0, 2, size=(64, 3)), dist=rng.uniform(0.1, 5, size=64), hit=1.0)
    irr, _ = O.ddgi_update(cfg, s)
    e = side + 1
    for t in range(1, side + 1):
        assert np.array_equal(irr[t, 0], irr[e - t, 1])        # left column: row mirrored
        assert np.array_equal(irr[t, e], irr[e - t, side])     # right column
        assert np.array_equal(irr[0, t], irr[1, e - t])        # top row: column mirrored
        assert np.array_equal(irr[e, t], irr[side, e - t])     # bottom row
    assert np.array_equal(irr[0, 0], irr[side, side]) and np.array_equal(irr[e, e], irr[1, 1])
    assert np.array_equal(irr[0, e], irr[side, 1]) and np.array_equal(irr[e, 0], irr[1, side])


def test_update_hysteresis_blend():
    cfg = ddgi_cfg(count=(1, 1, 1), rays=64, hysteresis=0.75)
    d = fib_dirs(64)
    a, _ = O.ddgi_update(cfg, surfels_from(d, [1.0, 2.0, 3.0], dist=1.0, hit=1.0))
    b, db = O.ddgi_update(cfg, surfels_from(d, [5.0, 6.0, 7.0], dist=3.0, hit=1.0), a, np.zeros(a.shape[:2] + (2,)))
    want = np.float32(0.75) * a[..., :3] + np.float32(0.25) * np.float32([5, 6, 7])
    np.testing.assert_allclose(b[..., :3], want, rtol=1e-6)
    np.testing.assert_allclose(db[..., 0], 0.25 * 3.0, rtol=1e-6)


def _uniform_field(cfg, irr_rgb, mean, m2):
    Hh, Ww = O.ddgi_atlas_shape(cfg)
    irr = np.zeros((Hh, Ww, 4), np.float32)
    irr[..., :3] = irr_rgb
    dep = np.zeros((Hh, Ww, 2), np.float32)
    dep[..., 0] = mean
    dep[..., 1] = m2
    return irr, dep


def test_lookup_uniform_field_and_energy():
    cfg = ddgi_cfg(count=(3, 2, 2), spacing=(1.0, 2.0, 0.5), energy=0.4)
    irr, dep = _uniform_field(cfg, [0.2, 1.0, 3.0], 100.0, 100.0 ** 2)
    rng = np.random.default_rng(5)
    n = rng.normal(size=(200, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    x = rng.uniform(-0.5, 2.5, size=(200, 3))
    v = -n
    out = O.ddgi_eval(cfg, irr, dep, np.concatenate([x, n, v], axis=1))
    np.testing.assert_allclose(out, np.broadcast_to(0.4 * np.float32([0.2, 1.0, 3.0]), out.shape), rtol=3e-6)


def _two_probe_blend(w0, w1, A, B, linear=False):
    """Blend of probes 0 and 1 with weights w0, w1 (floored at 1e-4); the six
    corners off the grid's single y and z layer keep the 1e-4 floor weight,
    three on each probe."""
    w0, w1 = max(1e-4, w0), max(1e-4, w1)
    if linear:
        return (w0 * A + w1 * B + 3e-4 * (A + B)) / (w0 + w1 + 6e-4)
    return ((w0 * math.sqrt(A) + w1 * math.sqrt(B) + 3e-4 * (math.sqrt(A) + math.sqrt(B))) / (w0 + w1 + 6e-4)) ** 2


def test_lookup_trilinear_between_two_probes():
    """Two probes on x at 0 and 1, both fully visible, normal +y, no bias:
    both wrap-shading weights are 1/2, the trilinear weights 1 - a and a, so
    the square-root blend is ((1 - a) sqrt(A) + a sqrt(B))^2 (linear: (1 - a)
    A + a B), up to the floor weight of the degenerate corners."""
    for flags in (0, L.RT_DDGI_LINEAR_BLENDING):
        cfg = ddgi_cfg(count=(2, 1, 1), spacing=(1.0, 1.0, 1.0), flags=flags)
        irr, dep = _uniform_field(cfg, 0.0, 1e4, 1e8)
        irr[:, :10, :3] = 0.5   # probe 0's tile
        irr[:, 10:, :3] = 2.0   # probe 1's tile
        for a in (0.0, 0.25, 0.5, 0.9):
            out = O.ddgi_eval(cfg, irr, dep, np.array([[a, 0, 0, 0, 1, 0, 0, -1, 0]], np.float32))
            want = _two_probe_blend(0.5 * (1 - a), 0.5 * a, 0.5, 2.0, linear=bool(flags))
            assert abs(out[0, 0] - want) < 1e-5 * want, (flags, a, out[0, 0], want)
            # the floor moves the textbook blend by less than 0.5%
            plain = (1 - a) * 0.5 + a * 2.0 if flags else ((1 - a) * math.sqrt(0.5) + a * math.sqrt(2.0)) ** 2
            assert abs(out[0, 0] - plain) < 5e-3 * plain


def test_lookup_chebyshev_visibility():
    """Probe 0 sees a wall at 0.1 in every direction (variance 0): the point
    half way to probe 1 (distance 0.5) is behind it, so its weight falls to
    the 1e-4 floor and the lookup is probe 1's value.  With the reference's
    condition (weight scaled only when distance <= mean) the occluded probe
    keeps its full weight."""
    for flags in (0, L.RT_DDGI_REFERENCE_VISIBILITY):
        cfg = ddgi_cfg(count=(2, 1, 1), spacing=(1.0, 1.0, 1.0), flags=flags)
        irr, dep = _uniform_field(cfg, 0.0, 1e4, 1e8)
        irr[:, :10, :3] = 0.5
        irr[:, 10:, :3] = 2.0
        dep[:, :10, 0] = 0.1
        dep[:, :10, 1] = 0.01
        out = O.ddgi_eval(cfg, irr, dep, np.array([[0.5, 0, 0, 0, 1, 0, 0, -1, 0]], np.float32))[0, 0]
        want = _two_probe_blend(0.25 if flags else 0.0, 0.25, 0.5, 2.0)
        assert abs(out - want) < 1e-5 * want, (flags, out, want)
    assert abs(_two_probe_blend(0.0, 0.25, 0.5, 2.0) - 2.0) < 0.01


def _sun(theta_deg=30.0, alpha_deg=5.0, S=40.0, spp=1, seed=3, max_depth=3, sky=0.0):
    t = math.radians(theta_deg)
    return make_settings(spp=spp, max_depth=max_depth, seed=seed, flags=0x4 | (0x2 if sky else 0),
                         sky=(sky, sky, sky), sun_dir=(math.sin(t), math.cos(t), 0.0), sun_radius_deg=alpha_deg,
                         sun_radiance=(S, S, S))


def test_probe_rays_over_sunlit_ground():
    """One probe 1 above an open ground (albedo rho) under sky K and a sun of
    half-angle alpha at zenith angle theta: upward rays see K (the sun is the
    light sample's), downward rays hit at 1 / |d_y| and carry the sun's
    closed form rho S sin^2(alpha) cos(theta) / 2 (the diffuse bounce's
    expectation in the path tracer's units)."""
    rho, S, K, th, al = 0.5, 40.0, 0.7, 30.0, 5.0
    sc = scenes.ground_plane((rho, rho, rho))
    cfg = ddgi_cfg(count=(1, 1, 1), start=(0.0, 1.0, 0.0), rays=256, shadow_count=16, max_dist=50.0)
    s = O.ddgi_surfels(sc, cfg, _sun(th, al, S, sky=K), np.eye(3), update=0)
    d = s[:, 0:3]
    np.testing.assert_allclose(d, fib_dirs(256), atol=0)
    up = d[:, 1] > 0
    assert np.all(s[up, 7] == 0) and np.all(s[up, 6] == 50.0)
    np.testing.assert_array_equal(s[up, 3:6], np.float32(K))
    dn = d[:, 1] < -0.05  # these reach the 200-wide ground within 20 units
    assert dn.sum() > 100 and np.all(s[dn, 7] == 1)
    np.testing.assert_allclose(s[dn, 6], 1.0 / np.abs(d[dn, 1]), rtol=2e-5)
    want = rho * S * math.sin(math.radians(al)) ** 2 * math.cos(math.radians(th)) / 2
    # the cone sample's value varies only through cos(theta) inside the cone
    np.testing.assert_allclose(s[dn, 3], want, rtol=0.1)
    assert abs(s[dn, 3].mean() - want) < 0.01 * want


def test_render_with_constant_field_over_sunlit_ground():
    """Constant field K with far depth: every ground pixel is the light
    sample's sun term plus pi rho K (albedo / 2 * lookup, times 2 pi):
    rho S pi sin^2(alpha) cos(theta) + pi rho K."""
    rho, S, K, th, al = 0.5, 40.0, 0.7, 30.0, 5.0
    sc = scenes.ground_plane((rho, rho, rho))
    cfg = ddgi_cfg(count=(2, 2, 2), start=(-2.0, 0.5, -2.0), spacing=(4.0, 1.0, 4.0))
    irr, dep = _uniform_field(cfg, K, 1e4, 1e8)
    cam = H.camera((0.0, 1.0, 0.0), 0.0, -1.0, math.pi / 4, 24, 24)
    rgb, nq = O.render_ddgi(sc, cam, _sun(th, al, S, spp=8, sky=K), cfg, irr, dep, q_vertex=0)
    assert nq == 24 * 24 * 8
    want = rho * S * math.pi * math.sin(math.radians(al)) ** 2 * math.cos(math.radians(th)) + math.pi * rho * K
    assert abs(rgb.mean() - want) < 3e-3 * want
    np.testing.assert_allclose(rgb, want, rtol=0.05)
