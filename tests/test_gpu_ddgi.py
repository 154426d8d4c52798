"""GPU irradiance probes (rt_ddgi_*): atlas update and lookup against the
oracle on identical inputs, one whole field update (probe rays, light
samples, multi-bounce from the previous field, random rotation) against the
oracle's, the render with the field against the oracle's render, a closed
form, and the handle's lifetime."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from master_thesis_amd import _lib as L
from master_thesis_amd import host_lib as H
from master_thesis_amd import scenes
from master_thesis_amd.probes import ProbeField
from master_thesis_amd.tracer import Tracer, make_settings

pytestmark = pytest.mark.gpu


def _field(tracer, **kw):
    args = dict(count=(3, 2, 3), start=(-2.0, -2.0, -2.0), spacing=(2.0, 4.0, 2.0), rays=64, side=6,
                max_distance=20.0, sharpness=2.0, normal_bias=0.05, shadow_count=4)
    args.update(kw)
    return ProbeField(tracer, **args)


def _random_surfels(n, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((n, 8), np.float32)
    d = rng.normal(size=(n, 3))
    s[:, 0:3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s[:, 3:6] = rng.uniform(0, 3, (n, 3))
    s[:, 6] = rng.uniform(0.05, 30, n)
    s[:, 7] = rng.integers(0, 2, n)
    return s


@pytest.mark.parametrize("sharpness", [0.0, 0.05, 2.0, 50.0])
def test_update_from_surfels_matches_oracle(tracer, sharpness):
    tracer.upload(scenes.cornell_box())
    f = _field(tracer, hysteresis=0.6, sharpness=sharpness)
    n = f.num_probes * f.rays
    s1, s2 = _random_surfels(n, 1), _random_surfels(n, 2)
    f.update_surfels(s1)
    g1 = f.atlas()
    f.update_surfels(s2)
    g2 = f.atlas()
    o1 = O.ddgi_update(f.cfg, s1)
    o2 = O.ddgi_update(f.cfg, s2, *o1)
    for g, o in ((g1, o1), (g2, o2)):
        # same sums in the same order: irradiance bit for bit.  The depth
        # weights are exp2(s log2 w) on the hardware's log and exp against
        # the oracle's powf: the weighted moments agree within 5e-7 relative
        # for sharpness 0.05 to 200 (exact at 0, where every weight is 1)
        np.testing.assert_array_equal(g[0], o[0])
        np.testing.assert_allclose(g[1], o[1], rtol=2e-6, atol=1e-6)
    f.close()


def test_lookup_matches_oracle(tracer):
    tracer.upload(scenes.cornell_box())
    for flags in ({}, dict(linear=True), dict(reference_visibility=True)):
        f = _field(tracer, **flags)
        Hh, Ww = f.atlas_shape
        rng = np.random.default_rng(7)
        irr = rng.uniform(0, 2, (Hh, Ww, 4)).astype(np.float32)
        mean = rng.uniform(0.2, 5, (Hh, Ww)).astype(np.float32)
        dep = np.stack([mean, mean * mean * rng.uniform(0.9, 1.3, (Hh, Ww))], axis=2).astype(np.float32)
        f.set_atlas(irr, dep)
        q = np.zeros((4096, 9), np.float32)
        q[:, 0:3] = rng.uniform(-3, 3, (4096, 3))
        nr = rng.normal(size=(4096, 3))
        q[:, 3:6] = nr / np.linalg.norm(nr, axis=1, keepdims=True)
        dv = rng.normal(size=(4096, 3))
        q[:, 6:9] = dv / np.linalg.norm(dv, axis=1, keepdims=True)
        got = f.eval(q)
        want = O.ddgi_eval(f.cfg, irr, dep, q)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-7)
        f.close()


def test_field_update_matches_oracle(tracer):
    """Two updates on the Cornell-style box (random rotation, multi-bounce):
    the second's surfels against the oracle's from the first's field."""
    sc = scenes.cornell_box()
    tracer.upload(sc)
    tracer.build()
    f = _field(tracer, rotate=True, multi_bounce=True, hysteresis=0.5)
    s = make_settings(spp=1, max_depth=4, seed=11)
    f.update(s)
    irr1, dep1 = f.atlas()
    f.update(s)
    got, rot = f.surfels()
    assert not np.allclose(rot, np.eye(3))
    np.testing.assert_allclose(rot @ rot.T, np.eye(3), atol=1e-6)
    want = O.ddgi_surfels(sc, f.cfg, s, rot, update=1, irr=irr1, dep=dep1, multi=True)
    np.testing.assert_array_equal(got[:, 0:3], want[:, 0:3])   # directions
    np.testing.assert_array_equal(got[:, 7], want[:, 7])       # hit flags
    np.testing.assert_array_equal(got[:, 6], want[:, 6])       # distances from bit-identical hits
    np.testing.assert_allclose(got[:, 3:6], want[:, 3:6], rtol=1e-5, atol=1e-6)
    assert (got[:, 3] > 0).mean() > 0.3
    irr2, dep2 = f.atlas()
    o = O.ddgi_update(f.cfg, got, irr1, dep1)
    np.testing.assert_array_equal(irr2, o[0])
    f.close()


def test_render_matches_oracle(tracer):
    sc = scenes.cornell_box()
    tracer.upload(sc)
    tracer.build()
    f = _field(tracer, multi_bounce=True, hysteresis=0.5)
    s = make_settings(spp=1, max_depth=4, seed=5)
    for _ in range(3):
        f.update(s)
    irr, dep = f.atlas()
    cam = H.camera(scenes.cornell_camera_position(), 1.0, 0.0, math.radians(40.0), 32, 32)
    rs = make_settings(spp=4, max_depth=5, seed=9)
    for qv in (0, 1):
        got, st = f.render(cam, rs, query_vertex=qv)
        want, nq = O.render_ddgi(sc, cam, rs, f.cfg, irr, dep, q_vertex=qv)
        assert st.queries == nq and nq > 0
        np.testing.assert_allclose(got, want[0], rtol=2e-4, atol=2e-5)
    f.close()


def test_render_constant_field_closed_form(tracer):
    """Open ground (albedo rho) under sky K and a sun: with a constant field K
    every pixel is rho S pi sin^2(alpha) cos(theta) + pi rho K."""
    rho, S, K, th, al = 0.5, 40.0, 0.7, 30.0, 5.0
    tracer.upload(scenes.ground_plane((rho, rho, rho)))
    tracer.build()
    f = _field(tracer, count=(2, 2, 2), start=(-2.0, 0.5, -2.0), spacing=(4.0, 1.0, 4.0))
    Hh, Ww = f.atlas_shape
    irr = np.zeros((Hh, Ww, 4), np.float32)
    irr[..., :3] = K
    dep = np.zeros((Hh, Ww, 2), np.float32)
    dep[..., 0] = 1e4
    dep[..., 1] = 1e8
    f.set_atlas(irr, dep)
    t = math.radians(th)
    s = make_settings(spp=16, max_depth=3, seed=3, flags=0x6, sky=(K, K, K), sun_dir=(math.sin(t), math.cos(t), 0.0),
                      sun_radius_deg=al, sun_radiance=(S, S, S))
    cam = H.camera((0.0, 1.0, 0.0), 0.0, -1.0, math.pi / 4, 48, 48)
    rgb, st = f.render(cam, s, query_vertex=0)
    want = rho * S * math.pi * math.sin(math.radians(al)) ** 2 * math.cos(t) + math.pi * rho * K
    assert st.queries == 48 * 48 * 16
    assert abs(rgb.mean() - want) < 2e-3 * want
    np.testing.assert_allclose(rgb, want, rtol=0.03)
    f.close()


def test_field_outliving_its_context():
    t = Tracer(0)
    t.upload(scenes.cornell_box())
    f = _field(t)
    lib, h = t.lib, f.h
    t.lib.rt_destroy(t.h)  # the field's device memory goes with the context
    t.h = None
    n = C.c_uint32()
    assert lib.rt_ddgi_atlas_size(h, C.byref(n), C.byref(n)) == L.RT_OK
    assert lib.rt_ddgi_update(h, C.byref(make_settings(1, 1))) == L.RT_E_STATE
    lib.rt_ddgi_destroy(h)
    f.h = None


def test_sw_render_cli_majercik2019(tracer, tmp_path):
    """The Vulkan renderer's Majercik2019 renderer file and a scene file with a
    Probes entry drive the GPU probes through the SoftwareRenderer-compatible
    command line: one field update, then shadowCountVision samples per pixel
    at the first hit; the same image as the library's (8-bit, up to the float
    sums' order)."""
    import subprocess
    from master_thesis_amd import build as B
    (tmp_path / "obj").mkdir()
    (tmp_path / "scene").mkdir()
    (tmp_path / "obj" / "panel.obj").write_text(
        "v 0.0 1.0 -1.0\nv 0.0 -1.0 -1.0\nv 0.0 1.0 1.0\nv 0.0 -1.0 1.0\nvn 1.0 0.0 0.0\nf 2//1 1//1 3//1 4//1\n")
    (tmp_path / "scene" / "box.scene").write_text(
        "panel.obj\n\tposition(-1, 0, 0)\n\n"
        "panel.obj\n\tposition(0, -1, 0)\n\trotation(0, 0, 1, 1.57079)\n\n"
        "panel.obj\n\tposition(0, 1, 0)\n\trotation(0, 0, 1, -1.57079)\n\n"
        "panel.obj\n\tposition(0, 0, -1)\n\trotation(0, 1, 0, -1.57079)\n\tcolor(0, 1, 0)\n\n"
        "panel.obj\n\tposition(0, 0, 1)\n\trotation(0, 1, 0, 1.57079)\n\tcolor(1, 0, 0)\n\n"
        "panel.obj\n\tposition(0, 0.99, 0)\n\tscale(0.2, 0.2, 0.2)\n\trotation(0, 0, 1, -1.57079)\n\tlightSource(1.0)\n"
        "\tcolor(15, 15, 15)\n\n"
        "Probes\n\tprobeCount(3, 3, 3)\n\tprobeStartCorner(-0.8, -0.8, -0.8)\n\tbetweenProbeDistance(0.8, 0.8, 0.8)\n")
    (tmp_path / "dd.renderer").write_text(
        "Majercik2019\n\tvisionJumpCount(3)\n\tperProbeRayCount(128)\n\tmaxProbeRayDistance(100)\n"
        "\tprobeSampleSideLength(6)\n\tdepthSharpness(4)\n\tnormalBias(0.02)\n\tlinearBlending(0)\n"
        "\tenergyPreservation(1)\n\tshadowCountProbe(2)\n\tshadowCountVision(8)\n\tseed(7)\n")
    (tmp_path / "c.camera").write_text("Camera\n\tposition(3.4, 0, 0)\n\tangle(1, 0)\n")
    out = tmp_path / "dd.ppm"
    p = subprocess.run([str(B.CLI), str(tmp_path / "dd.renderer"), str(tmp_path / "scene" / "box.scene"), "64", "48",
                        "4", str(tmp_path / "c.camera"), str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    head = b"P6\n64\n48\n255\n"
    img = np.frombuffer(out.read_bytes()[len(head):], np.uint8).reshape(48, 64, 3).astype(int)
    sc = H.load_scene(str(tmp_path / "scene" / "box.scene"), str(tmp_path / "obj"))
    tracer.upload(sc)
    tracer.build()
    f = ProbeField(tracer, (3, 3, 3), (-0.8, -0.8, -0.8), (0.8, 0.8, 0.8), rays=128, side=6, max_distance=100.0,
                   sharpness=4.0, normal_bias=0.02, energy=1.0, shadow_count=2)
    f.update(make_settings(spp=8, max_depth=3, seed=7))
    cam = H.camera_from_file(str(tmp_path / "c.camera"), 64, 48)
    rgb, st = f.render(cam, make_settings(spp=8, max_depth=3, seed=7), query_vertex=0)
    f.close()
    want = (np.clip(rgb, 0.0, 1.0) * np.float32(255.0)).astype(int)
    d = np.abs(img - want)
    assert d.max() <= 1 and (d > 0).mean() < 0.02
    assert st.queries > 0 and img.mean() > 5
    # a scene without a Probes entry is an error for this renderer
    (tmp_path / "scene" / "np.scene").write_text("panel.obj\n\tposition(-1, 0, 0)\n")
    p = subprocess.run([str(B.CLI), str(tmp_path / "dd.renderer"), str(tmp_path / "scene" / "np.scene"), "8", "8",
                        "1", str(tmp_path / "c.camera"), str(tmp_path / "x.ppm")], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 1 and "no Probes entry" in p.stderr
