"""Oracle and host layer against the reference itself, compiled from its own
sources (oracle/_ref/libref.so).  Runs where the reference was available at
build time; the GPU box checks the golden vectors made from it instead
(tests/golden/)."""
import glob
import math
import os

import numpy as np
import pytest

import oracle as O
from oracle import ref as R
from master_thesis_amd import scenes, host_lib as H
from master_thesis_amd._lib import RAY_DTYPE, RT_MISS
from master_thesis_amd.tracer import make_settings

REF_RES = "/root/reference/res"
pytestmark = pytest.mark.skipif(not R.available() or not os.path.isdir(REF_RES),
                                reason="compiled reference (oracle/_ref) or its resources not present here")


def _random_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY_DTYPE)
    r["o"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    r["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["tmax"] = np.inf
    return r


def test_moller_trumbore_matches_reference():
    rng = np.random.default_rng(2)
    hits = 0
    for _ in range(4000):
        v = rng.normal(size=(3, 3)).astype(np.float32)
        o = rng.normal(size=3).astype(np.float32) * 3
        target = (v.mean(axis=0) + rng.normal(size=3) * 0.7).astype(np.float32)
        d = (target - o).astype(np.float32)
        rh, rp, rb = R.intersect(v[0], v[1], v[2], o, d)
        oh, tuv = O.intersect(o, d, v[0], v[1], v[2])
        assert rh == oh
        if oh:
            hits += 1
            p = (o + tuv[0] * d).astype(np.float32)
            assert np.array_equal(p.view(np.uint32), rp.view(np.uint32))
            assert np.array_equal(O.barycentric(v[0], v[1], v[2], p).view(np.uint32), rb.view(np.uint32))
    assert hits > 500


def test_scene_trace_matches_reference_up_to_exact_ties():
    sc = scenes.cornell_box()
    rays = _random_rays(20000, -2.7, 2.7, 1)
    rh, rp, rn, robj = R.RefScene(sc).trace(rays)
    hb = O.trace_brute(sc, rays)
    oh = hb["tri"] != RT_MISS
    assert np.array_equal(rh, oh)
    p, nr = O.hit_shading(sc, rays, hb)
    diff = oh & (np.any(p != rp, axis=1) | np.any(nr != rn, axis=1) | (sc.tri_mat[np.where(oh, hb["tri"], 0)] != robj))
    # every difference must be an exact squared-distance tie between two triangles
    for i in np.nonzero(diff)[0]:
        d2 = []
        for t in range(sc.idx.shape[0]):
            v = sc.pos[sc.idx[t]]
            ok, tuv = O.intersect(rays["o"][i], rays["d"][i], v[0], v[1], v[2])
            if ok:
                w = (rays["o"][i] - (rays["o"][i] + tuv[0] * rays["d"][i]).astype(np.float32)).astype(np.float32)
                d2.append(np.float32(np.float32(np.float32(np.float32(0) + w[0] * w[0]) + w[1] * w[1]) + w[2] * w[2]))
        d2 = np.sort(np.array(d2, np.float32))
        assert len(d2) >= 2 and d2[0] == d2[1], f"ray {i}: not a tie"
    assert diff.mean() < 0.01  # the boxes stand on the floor: their bottom faces tie with it


@pytest.mark.parametrize("pos,a0,a1", [((3.4, 0, 0), 1, 0), ((4, 0, 0), 1, 0),
                                       ((8.29397, 0, -1.6169), 0.783549, -0.0445634), ((0, 60, 0), 1, -1),
                                       ((1, 2, 3), 1.7, 0.3), ((-2, 0.5, 1), -1.3, 1.4), ((0.1, 0.2, 0.3), 0.0, -0.7)])
def test_camera_matches_reference(pos, a0, a1):
    vi, pi, og = R.camera(pos, a0, a1, 320, 200)
    cam = H.camera(pos, a0, a1, np.float32(math.pi / 4), 320, 200)
    assert np.array_equal(vi.view(np.uint32), np.array(cam.view_inv[:], np.float32).view(np.uint32))
    assert np.array_equal(pi.view(np.uint32), np.array(cam.proj_inv[:], np.float32).view(np.uint32))
    assert np.array_equal(og.view(np.uint32), np.array(cam.origin[:], np.float32).view(np.uint32))


@pytest.mark.parametrize("scene", sorted(glob.glob(f"{REF_RES}/scene/*.scene")))
def test_scene_files_flatten_like_reference(scene):
    pos, nrm, idx, tri_obj, colors = R.load_scene(scene, "/root/reference/build/")
    mine = H.load_scene(scene, f"{REF_RES}/obj")
    assert np.array_equal(mine.pos.view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(mine.nrm.view(np.uint32), nrm.view(np.uint32))
    assert np.array_equal(mine.idx, idx)
    assert np.array_equal(mine.tri_mat, tri_obj)
    assert np.array_equal(mine.materials["color"], colors)


def test_path_tracer_statistically_matches_reference():
    """The reference's own PathTracer (mt19937 stream) and the oracle
    (Philox stream) estimate the same image: block means agree within the
    spread the oracle shows across seeds.  The reference seeds its streams
    afresh on every call, so its image differs run to run; the spread is taken
    over 16 oracle seeds, which keeps the t statistic's tails (15 degrees of
    freedom) far below the threshold for an unbiased estimator while a dropped
    term or a wrong weight moves block means by tens of sigma."""
    sc = scenes.cornell_box()
    cam = H.camera(scenes.cornell_camera_position(), 1.0, 0.0, np.float32(math.pi / 4), 32, 32)
    spp, depth = 256, 5
    ref = R.RefScene(sc).render(cam, spp, depth, threads=8)
    runs = np.stack([O.render(sc, cam, make_settings(spp=spp, max_depth=depth, seed=s))[0] for s in range(1, 17)])

    def blocks(img):
        return img.reshape(4, 8, 4, 8, 3).mean(axis=(1, 3))

    rb = blocks(ref)
    ob = np.stack([blocks(x) for x in runs])
    m, s = ob.mean(axis=0), ob.std(axis=0, ddof=1)
    z = (rb - m) / np.maximum(s * math.sqrt(1 + 1 / len(runs)), 1e-6)
    assert np.all(np.abs(z) < 7), np.abs(z).max()
    zg = (ref.mean() - runs.mean()) / (runs.mean(axis=(1, 2, 3)).std(ddof=1) * math.sqrt(1 + 1 / len(runs)))
    assert abs(zg) < 5, zg


def test_occluded_matches_reference():
    """Scene::isOccluded of the compiled reference against the oracle's brute
    force on fresh segments, some ending exactly at hit points."""
    sc = scenes.cornell_box()
    ref = R.RefScene(sc)
    rng = np.random.default_rng(77)
    n = 3000
    a = rng.uniform(-2.7, 2.7, (n, 3)).astype(np.float32)
    b = rng.uniform(-2.7, 2.7, (n, 3)).astype(np.float32)
    rays = _random_rays(n, -2.7, 2.7, 78)
    hit, p, _, _ = ref.trace(rays)
    segs = np.concatenate([np.concatenate([a, b], axis=1),
                           np.concatenate([rays["o"][hit], p[hit]], axis=1)]).astype(np.float32)
    np.testing.assert_array_equal(O.occluded_brute(sc, segs), ref.occluded(segs))


def test_bidirectional_tracer_statistically_matches_reference():
    """The reference's own BidirectionalPathTracer (its mt19937 stream) and
    the oracle's restatement (Philox) estimate the same image: block means
    agree within the spread the oracle shows across seeds."""
    sc = scenes.cornell_box()
    cam = H.camera(scenes.cornell_camera_position(), 1.0, 0.0, np.float32(math.pi / 4), 32, 32)
    spp = 64
    ref = R.RefScene(sc).render_bdpt(cam, spp, 5, 5, 5, threads=8)
    runs = np.stack([O.render_bdpt(sc, cam, make_settings(spp=spp, max_depth=5, seed=s)) for s in range(1, 17)])

    def blocks(img):
        return img.reshape(4, 8, 4, 8, 3).mean(axis=(1, 3))

    rb = blocks(ref)
    ob = np.stack([blocks(x) for x in runs])
    m, s = ob.mean(axis=0), ob.std(axis=0, ddof=1)
    z = (rb - m) / np.maximum(s * math.sqrt(1 + 1 / len(runs)), 1e-6)
    assert np.all(np.abs(z) < 7), np.abs(z).max()
    zg = (ref.mean() - runs.mean()) / (runs.mean(axis=(1, 2, 3)).std(ddof=1) * math.sqrt(1 + 1 / len(runs)))
    assert abs(zg) < 5, zg


def test_reflect_refract_bit_exact_against_reference():
    """reflect and customRefract (math/vector.h:319-345) of the oracle against
    the reference's own templates, bit for bit: random unit and non-unit
    directions on both sides of the normal, indices 0.5-3 (both branches of
    customRefract, total internal reflection included), normal and grazing
    incidence."""
    rng = np.random.default_rng(11)
    kinds = {"refract": 0, "reflect": 0}
    cases = []
    for k in range(40000):
        N = rng.normal(size=3)
        N /= np.linalg.norm(N)
        I = rng.normal(size=3)
        if k % 4 != 3:
            I /= np.linalg.norm(I)
        if k % 50 == 0:
            I = -N + rng.normal(size=3) * 1e-4          # near-normal incidence
        if k % 50 == 1:
            I = np.cross(N, rng.normal(size=3))         # grazing
            I = I / np.linalg.norm(I) + N * 1e-3
        cases.append((I.astype(np.float32), N.astype(np.float32), np.float32(rng.uniform(0.5, 3.0))))
    for I, N, eta in cases:
        a, b = O.custom_refract(I, N, eta), R.custom_refract(I, N, eta)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (I, N, eta, a, b)
        ra, rb = O.reflect(I, N), R.reflect(I, N)
        assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32)), (I, N, ra, rb)
        kinds["reflect" if np.array_equal(a, ra) else "refract"] += 1
    assert kinds["reflect"] > 2000 and kinds["refract"] > 20000, kinds


def _blocks_match(ref, runs, nb=4):
    h, w = ref.shape[:2]

    def blocks(img):
        return img.reshape(nb, h // nb, nb, w // nb, 3).mean(axis=(1, 3))

    rb = blocks(ref)
    ob = np.stack([blocks(x) for x in runs])
    m, s = ob.mean(axis=0), ob.std(axis=0, ddof=1)
    z = (rb - m) / np.maximum(s * math.sqrt(1 + 1 / len(runs)), 1e-6)
    assert np.all(np.abs(z) < 7), np.abs(z).max()
    zg = (ref.mean() - runs.mean()) / (runs.mean(axis=(1, 2, 3)).std(ddof=1) * math.sqrt(1 + 1 / len(runs)))
    assert abs(zg) < 5, zg


def test_path_tracer_specular_statistically_matches_reference():
    """The specular and transmissive branches (path_tracer.cpp:62-71) end to
    end: the Cornell-style box with a mirror box, a glass ball and a ball
    mixing all three branches, rendered by the reference's PathTracer and by
    the oracle, agree block by block within the spread across oracle seeds.
    Weights and indices go to the reference objects, which derive their own
    thresholds."""
    sc = scenes.cornell_specular()
    cam = H.camera(scenes.cornell_camera_position(), 1.0, 0.0, np.float32(math.pi / 4), 32, 32)
    spp, depth = 256, 6
    ref = R.RefScene(sc, scenes.SPECULAR_WEIGHTS).render(cam, spp, depth, threads=8)
    runs = np.stack([O.render(sc, cam, make_settings(spp=spp, max_depth=depth, seed=s))[0] for s in range(1, 17)])
    _blocks_match(ref, runs)
    # the same test tells the specular scene from one whose mirror and balls
    # are diffuse (the branches dropped): some blocks move by many sigma
    flat = sc.materials.copy()
    for k in range(4, len(flat)):
        flat[k] = scenes.material(flat[k]["color"])
    diffuse = scenes.Scene(sc.pos, sc.nrm, sc.idx, sc.tri_mat, flat)
    dimg = O.render(diffuse, cam, make_settings(spp=spp, max_depth=depth, seed=99))[0]
    with pytest.raises(AssertionError):
        _blocks_match(dimg, runs)


def test_ball_scene_statistically_matches_reference():
    """The reference's own res/scene/cornell_box_with_ball.scene (a glass ball,
    transparent(1), refractionIndex(1.5)) read by the reference's MeshManager
    and rendered by its PathTracer, against the same file read by the host
    layer and rendered by the oracle."""
    path = f"{REF_RES}/scene/cornell_box_with_ball.scene"
    sc = H.load_scene(path, f"{REF_RES}/obj")
    m = sc.materials[sc.tri_mat[-1]]
    assert m["transparent_threshold"] == 1.0 and m["diffuse_threshold"] == 0.0 and m["refraction_index"] == np.float32(1.5)
    sc.materials["color"][sc.materials["light_source"] == 1] = 15.0
    ref_scene = R.RefScene(sc, np.stack([np.array([1, 0, 0, 1], np.float32)] * (len(sc.materials) - 1)
                                        + [np.array([0, 0, 1, 1.5], np.float32)]))
    cam = H.camera((3.4, 0.0, 0.0), 1.0, 0.0, np.float32(math.pi / 4), 32, 32)
    spp, depth = 256, 6
    ref = ref_scene.render(cam, spp, depth, threads=8)
    runs = np.stack([O.render(sc, cam, make_settings(spp=spp, max_depth=depth, seed=s))[0] for s in range(1, 17)])
    _blocks_match(ref, runs)
    # the same file through the reference's own loader gives the same scene
    rf = R.RefScene.from_file(path, "/root/reference/build/")
    rays = _random_rays(4000, -0.95, 0.95, 12)
    a, b = rf.trace(rays), ref_scene.trace(rays)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_bidirectional_tracer_specular_statistically_matches_reference():
    """The reference's BidirectionalPathTracer on the box with the mirror and
    the glass and mixed balls (its vision and light paths take the specular
    branches, bidirectional_path_tracer.cpp:104-110) against the oracle."""
    sc = scenes.cornell_specular()
    cam = H.camera(scenes.cornell_camera_position(), 1.0, 0.0, np.float32(math.pi / 4), 32, 32)
    spp = 64
    ref = R.RefScene(sc, scenes.SPECULAR_WEIGHTS).render_bdpt(cam, spp, 6, 5, 6, threads=8)
    runs = np.stack([O.render_bdpt(sc, cam, make_settings(spp=spp, max_depth=6, seed=s), 6, 5, 6)
                     for s in range(1, 17)])
    _blocks_match(ref, runs)


def test_faithful_mode_statistically_matches_reference():
    """The CPU baseline's faithful mode (oracle/faithful.cpp: the reference's
    BVHs, queries, shared mutex-guarded mt19937 and pixel counter restated)
    renders what the compiled reference renders, objects grouped alike (one
    per quad or box): block means within the spread across its own seeds."""
    sc = scenes.cornell_box()
    tri_obj, obj_mat = sc.objects()
    cam = H.camera(scenes.cornell_camera_position(), 1.0, 0.0, np.float32(math.pi / 4), 32, 32)
    spp, depth = 256, 5
    obj_scene = scenes.Scene(sc.pos, sc.nrm, sc.idx, tri_obj, sc.materials[obj_mat])
    ref = R.RefScene(obj_scene).render(cam, spp, depth, threads=8)
    runs = np.stack([O.render_faithful(sc, cam, spp, depth, threads=1, seed=s)[0] for s in range(1, 17)])
    _blocks_match(ref, runs)
