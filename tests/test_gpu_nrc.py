"""GPU radiance cache: MFMA inference and gradient against the CPU
reference network with the same BF16 roundings (oracle/nrc.py), parameter
round trip, online training that lowers the loss; path tracing with the
cache against the oracle's constant-cache render, a closed form, the plain
render, and a converged reference (the cache's gain)."""
import math

import numpy as np
import pytest

import oracle as O
from oracle import nrc as N
from master_thesis_amd import host_lib as H
from master_thesis_amd import scenes
from master_thesis_amd.cache import RadianceCache
from master_thesis_amd.tracer import make_settings

pytestmark = pytest.mark.gpu

LO = np.array([-1.0, -1.0, -1.0], np.float32)
HI = np.array([1.0, 1.0, 1.0], np.float32)
INV = (1.0 / (HI - LO)).astype(np.float32)


def _samples(n, seed=0):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 9), np.float32)
    f[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    nr = rng.normal(size=(n, 3))
    f[:, 3:6] = nr / np.linalg.norm(nr, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    f[:, 6:9] = d / np.linalg.norm(d, axis=1, keepdims=True)
    return f


def _target(f):
    """A smooth radiance-like field of position and normal."""
    p, nrm = f[:, 0:3], f[:, 3:6]
    base = 0.6 + 0.4 * np.sin(2.0 * p[:, 0:1] + np.array([0.0, 1.0, 2.0])) * np.cos(1.5 * p[:, 1:2])
    return (base * (0.75 + 0.25 * nrm[:, 2:3])).astype(np.float32)


@pytest.fixture(scope="module")
def cache(tracer):
    tracer.upload(scenes.cornell_box())
    c = RadianceCache(tracer, lr=1e-2, seed=3, bounds=(LO, HI))
    yield c
    c.close()


def _set_random(cache, seed):
    p = N.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    for l in range(6):  # small non-zero biases so every bias path is exercised
        nv = 3 if l == 5 else 64
        p[N.OFFB[l]:N.OFFB[l] + nv] = rng.normal(0, 0.05, nv)
    cache.set_params(p)
    return p


def test_params_round_trip(cache):
    assert cache.num_params() == N.NPARAM
    p = _set_random(cache, 11)
    np.testing.assert_array_equal(cache.params(), p)


# Tolerances: the reference rounds inputs, weights, activations and
# back-propagated gradients to BF16 exactly where the kernels do and sums in
# float64; the kernels sum in FP32 MFMA order and encode with the hardware
# sine and cosine of three anchors and angle doubling.
# Either can flip a BF16 rounding (2^-8 relative) now and then, so the bar is
# well below the BF16 quantisation error itself (emulated vs exact network).
@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4097])
def test_infer_matches_reference(cache, n):
    p = _set_random(cache, 21)
    f = _samples(n, n)
    got = cache.infer(f)
    want, _ = N.forward(p, f, LO, INV, emulate=True)
    exact, _ = N.forward(p, f, LO, INV, emulate=False)
    err = np.abs(got - want)
    q = np.abs(want - exact)
    scale = max(1.0, float(np.abs(exact).max()))
    assert err.max() <= 2e-2 * scale, err.max()
    assert np.sqrt(np.mean(err ** 2)) <= max(0.25 * np.sqrt(np.mean(q ** 2)), 1e-6 * scale)


def test_infer_device_pointers(cache):
    import torch
    p = _set_random(cache, 22)
    f = _samples(5000, 7)
    host = cache.infer(f)
    dev = cache.infer(torch.from_numpy(f).cuda())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy(), host)


# 64 .. 4096: the 4-wave training kernel; 40000: the 8-wave one (tiles of
# 128 samples fill every CU), its last tile ragged
@pytest.mark.parametrize("n", [64, 777, 4096, 40000])
def test_gradient_matches_reference(cache, n):
    p = _set_random(cache, 31)
    f = _samples(n, 100 + n)
    t = _target(f)
    g, loss = cache.gradient(f, t)
    loss_ref, g_ref = N.loss_and_grad(p, f, t, LO, INV, emulate=True)
    _, g_exact = N.loss_and_grad(p, f, t, LO, INV, emulate=False)
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    # per layer: relative L2 error of the gradient block well below the BF16 error
    for l in range(6):
        sl = slice(N.OFFW[l], N.OFFB[l] + N.ND[l])
        ref = g_ref[sl]
        nrm = np.linalg.norm(ref)
        e = np.linalg.norm(g[sl] - ref) / nrm
        q = np.linalg.norm(g_exact[sl] - ref) / nrm
        assert e <= max(0.5 * q, 2e-3), (l, e, q)
    # padded weights get exactly zero gradient
    assert np.all(g[:N.ND[0] * N.KD[0]].reshape(64, 96)[:, 78:] == 0.0)
    assert np.all(g[N.OFFW[5]:N.OFFB[5]].reshape(16, 64)[3:] == 0.0)


@pytest.mark.parametrize("n", [3000, 40000])
def test_gradient_is_deterministic(cache, n):
    _set_random(cache, 41)
    f = _samples(n, 5)
    t = _target(f)
    g1, l1 = cache.gradient(f, t)
    g2, l2 = cache.gradient(f, t)
    np.testing.assert_array_equal(g1, g2)
    assert l1 == l2


def test_training_lowers_loss(cache):
    _set_random(cache, 51)
    f_eval = _samples(4096, 999)
    t_eval = _target(f_eval)
    _, loss0 = cache.gradient(f_eval, t_eval)
    for step in range(200):
        f = _samples(4096, 1000 + step)
        cache.train(f, _target(f), want_loss=False)
    _, loss1 = cache.gradient(f_eval, t_eval)
    assert loss1 < 0.1 * loss0, (loss0, loss1)
    y = cache.infer(f_eval)
    rel = np.sqrt(np.mean((y - t_eval) ** 2)) / np.sqrt(np.mean(t_eval ** 2))
    assert rel < 0.15, rel


# ---- path tracing with the cache (rt_render_nrc)


def _constant_params(c):
    """All weights zero, output bias c: the network predicts c everywhere."""
    p = np.zeros(N.NPARAM, np.float32)
    p[N.OFFB[5]:N.OFFB[5] + 3] = c
    return p


def _scene_cache(tracer, sc, seed=0, cluster=False):
    tracer.upload(sc)
    tracer.build(2, cluster=cluster)
    return RadianceCache(tracer, lr=1e-2, seed=seed)


def test_render_without_queries_is_the_plain_render(tracer):
    sc = scenes.cornell_box()
    cache = _scene_cache(tracer, sc)
    cam = H.cornell_camera(48, 40)
    s = make_settings(spp=8, max_depth=5, seed=3)
    plain, _ = tracer.render(cam, s)
    got, st = cache.render(cam, s, train=False, query=False, nee=False)
    assert st.queries == 0 and st.records > 0 and st.steps == 0  # records are made, no step taken
    np.testing.assert_array_equal(got, plain)
    cache.close()


@pytest.mark.parametrize("q_vertex,every,nee", [(1, 16, False), (2, 4, False), (1, 0, False),
                                                 (1, 16, True), (2, 4, True), (1, 0, True)])
def test_constant_cache_render_matches_oracle(tracer, q_vertex, every, nee):
    """Same queries, records and image as the oracle's constant-cache render
    (images to float summation order: the GPU adds cache terms and light
    samples atomically)."""
    sc = scenes.cornell_box()
    cache = _scene_cache(tracer, sc)
    c = np.array([0.4, 0.25, 0.1], np.float32)
    cache.set_params(_constant_params(c))
    cam = H.cornell_camera(40, 32)
    s = make_settings(spp=8, max_depth=6, seed=9, sample_offset=5)
    got, st = cache.render(cam, s, query_vertex=q_vertex, train_every=every, train=False, nee=nee)
    want, nq, nr, wf, wt, _ = O.render_cached(sc, cam, s, c, q_vertex, every, nee=nee, records=True)
    assert (st.queries, st.records) == (nq, nr)
    assert nq > 0 and (st.shadow_rays > 0) == nee
    np.testing.assert_allclose(got, want[0], rtol=1e-4, atol=1e-5)
    # the training records, in (view, pixel, sample) order on both sides:
    # same features, targets summed in the same order of path vertices
    gf, gt = cache.last_records()
    assert gf.shape[0] == nr
    if nr:
        np.testing.assert_array_equal(gf, wf)
        np.testing.assert_array_equal(gt, wt)
    cache.close()


def test_sun_sampling_matches_closed_form(tracer):
    rho, S, th, al = 0.5, 40.0, 30.0, 5.0
    sc = scenes.ground_plane((rho, rho, rho))
    cache = _scene_cache(tracer, sc)
    cam = H.camera((0.0, 1.0, 0.0), 0.0, -1.0, math.pi / 4, 64, 64)
    t = math.radians(th)
    s = make_settings(spp=8, max_depth=3, seed=3, flags=0x4, sun_dir=(math.sin(t), math.cos(t), 0.0),
                      sun_radius_deg=al, sun_radiance=(S, S, S))
    got, st = cache.render(cam, s, train_every=0, train=False, query=False)
    want = rho * S * math.pi * math.sin(math.radians(al)) ** 2 * math.cos(t)
    assert st.shadow_rays == 64 * 64 * 8
    assert abs(got.mean() - want) < 2e-3 * want
    wo, _, _ = O.render_cached(sc, cam, s, np.zeros(3), 1, 0, query=False, nee=True)
    np.testing.assert_allclose(got, wo[0], rtol=1e-5)
    cache.close()


def test_constant_cache_closed_room_is_pi_rho_c(tracer):
    rho = 0.6
    sc = scenes.closed_room(scenes.material((rho, rho, rho)), scenes.material(scenes.WHITE))
    cache = _scene_cache(tracer, sc)
    c = np.array([0.3, 0.5, 0.7], np.float32)
    cache.set_params(_constant_params(c))
    cam = H.camera((0.0, 0.0, 0.0), 0.0, -1.0, math.pi / 4, 64, 64)
    N_ = 16
    got, st = cache.render(cam, make_settings(spp=N_, max_depth=4, seed=2), train_every=0, train=False)
    P = 64 * 64
    assert st.queries == N_ * P
    for ch in range(3):
        sd = 2 * math.pi * rho * c[ch] / math.sqrt(12 * N_ * P)
        assert abs(got[..., ch].mean() - math.pi * rho * c[ch]) < 4 * sd
    cache.close()


def test_online_training_is_deterministic(tracer):
    """Training records are ordered by their samples before the batches are
    drawn, and each record's target has one writer at a time: two caches from
    the same initial weights trained on the same frames end with the same
    parameters, bit for bit, and render the same records."""
    sc = scenes.cornell_closed()
    cam = H.closed_camera(96, 96)
    out = []
    for run in range(2):
        cache = _scene_cache(tracer, sc, seed=4)
        for f in range(8):
            _, st = cache.render(cam, make_settings(spp=1, max_depth=8, seed=7, sample_offset=f), train_every=4,
                                 batch=512)
            assert st.steps == 4 and st.records > 2048
        out.append((cache.params(), *cache.last_records(), st.loss))
        cache.close()
    (p0, f0, t0, l0), (p1, f1, t1, l1) = out
    np.testing.assert_array_equal(f0, f1)
    np.testing.assert_array_equal(t0, t1)
    np.testing.assert_array_equal(p0, p1)
    assert l0 == l1


def test_online_training_gain_and_bias(tracer):
    """Closed Cornell-style box (no path escapes), 256 x 256, depth 8.
    Reference: 8192 light-sampled spp, whose mean the plain render confirms
    (the plain estimator is the yardstick).  After 200 frames of online training
    at 1 spp, 4-spp frames that stop at the cache after the first bounce reach
    the RMSE of light-sampled frames without the network while tracing less
    than half their rays (extension plus shadow rays), and the cache's bias
    (the cached render at 2048 spp against the reference) stays within 3% of
    the frame mean.  Training is deterministic (records in sample order), so
    these figures are those of one fixed network: cached / light-sampled
    RMSE 0.92, a quarter of the rays, bias -1.5% of the frame mean.  The bias
    is a property of the trained network: five training sequences (network
    and path seeds) gave -1.6% .. +0.6%, and a change of summation order in
    the trainer moves it within that spread."""
    sc = scenes.cornell_closed()
    cache = _scene_cache(tracer, sc, seed=1)
    cam = H.closed_camera(256, 256)
    ref, _ = cache.render(cam, make_settings(spp=8192, max_depth=8, seed=100), train=False, query=False,
                          train_every=0)
    plain_ref, _ = tracer.render(cam, make_settings(spp=8192, max_depth=8, seed=101))
    assert abs(ref.mean() - plain_ref.mean()) < 0.005 * ref.mean()
    for f in range(200):
        cache.render(cam, make_settings(spp=1, max_depth=8, seed=7, sample_offset=f), train_every=4, batch=4096)
    err = {"plain": [], "light": [], "cached": []}
    rays = {"light": 0, "cached": 0}
    for f in range(4):
        s = make_settings(spp=4, max_depth=8, seed=7, sample_offset=100000 + 4 * f)
        err["plain"].append(np.sqrt(np.mean((tracer.render(cam, s)[0] - ref) ** 2)))
        img, st = cache.render(cam, s, train=False, query=False, train_every=0)
        err["light"].append(np.sqrt(np.mean((img - ref) ** 2)))
        rays["light"] += tracer.timing()["rays_traced"] + st.shadow_rays
        img, st = cache.render(cam, s, train=False, train_every=16)
        err["cached"].append(np.sqrt(np.mean((img - ref) ** 2)))
        rays["cached"] += tracer.timing()["rays_traced"] + st.shadow_rays
    e = {k: float(np.mean(v)) for k, v in err.items()}
    big, _ = cache.render(cam, make_settings(spp=2048, max_depth=8, seed=7, sample_offset=200000), train=False,
                          train_every=0)
    bias = float((big - ref).mean())
    print("rmse at 4 spp:", e, "rays:", rays, "cache bias (frame mean):", bias, "relative:", bias / ref.mean())
    assert e["light"] < 0.1 * e["plain"]
    assert e["cached"] < 1.05 * e["light"]
    assert rays["cached"] < 0.5 * rays["light"]
    # the cache's bias after 800 Adam steps
    assert abs(bias) < 0.03 * ref.mean()
    cache.close()


# ---- the cache on its own workload: the 9.69M-triangle terrain, sun and sky

def _terrain_settings(spp, off=0, seed=5):
    from master_thesis_amd._lib import RT_RENDER_SKY, RT_RENDER_SUN
    return make_settings(spp=spp, max_depth=8, seed=seed, flags=RT_RENDER_SKY | RT_RENDER_SUN, sample_offset=off,
                         sky=(0.45, 0.55, 0.75), sun_dir=(0.4, 0.8, 0.3), sun_radius_deg=5.0,
                         sun_radiance=(40.0, 36.0, 30.0))


@pytest.fixture(scope="module")
def terrain():
    return scenes.terrain_scene()


def test_terrain_constant_cache_matches_oracle(tracer, terrain):
    """rt_render_nrc on the full terrain (the learned cache's workload, 2048^2
    heightfield + 64 icospheres): two 48 x 27 views, 2 spp, depth 8, light
    sampling, a constant cache: queries and records as the oracle's
    constant-cache render, which builds its own BVH; records in the same
    order with the same features and targets; images to summation order."""
    cache = _scene_cache(tracer, terrain, cluster=True)  # the terrain workloads' hierarchy
    c = np.array([0.5, 0.4, 0.3], np.float32)
    cache.set_params(_constant_params(c))
    cams = scenes.terrain_cameras(2, 48, 27)
    s = _terrain_settings(2, off=3)
    got, st = cache.render(cams, s, query_vertex=1, train_every=4, train=False)
    want, nq, nr, wf, wt, _ = O.render_cached(terrain, cams, s, c, 1, 4, nee=True, records=True, threads=16)
    assert (st.queries, st.records) == (nq, nr) and nq > 0 and nr > 0 and st.shadow_rays > 0
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-5)
    gf, gt = cache.last_records()
    np.testing.assert_array_equal(gf, wf)
    np.testing.assert_array_equal(gt, wt)
    assert got.mean() > 0.1
    cache.close()


def test_terrain_online_training_gain(tracer, terrain):
    """Online training on the terrain at 128 x 72, 1 spp per frame: after 60
    frames the cached 4-spp frames are closer to a light-sampled reference
    (1024 spp) than those of the untrained network, reach the RMSE of
    light-sampled frames without the network while tracing fewer rays, and
    the cache's bias (its 256-spp image against the reference) stays within
    a few percent of the frame mean."""
    cam = scenes.terrain_cameras(1, 128, 72)[0]
    cache = _scene_cache(tracer, terrain, seed=1, cluster=True)
    ref, _ = cache.render(cam, _terrain_settings(1024, 0, 99), train=False, query=False, train_every=0)

    def frames(mode):
        err, rays = [], 0
        for f in range(4):
            s = _terrain_settings(4, 500000 + 4 * f)
            if mode == "light":
                img, st = cache.render(cam, s, train=False, query=False, train_every=0)
            else:
                img, st = cache.render(cam, s, train=False, train_every=0)
            err.append(float(np.sqrt(np.mean((img - ref) ** 2))))
            rays += tracer.timing()["rays_traced"] + st.shadow_rays
        return float(np.mean(err)), rays

    e_untrained, _ = frames("cached")
    losses = []
    for f in range(60):
        _, st = cache.render(cam, _terrain_settings(1, f, 7), train_every=1, batch=2048)
        losses.append(st.loss)
    e_light, r_light = frames("light")
    e_cached, r_cached = frames("cached")
    big, _ = cache.render(cam, _terrain_settings(256, 700000), train=False, train_every=0)
    bias = float((big - ref).mean()) / float(ref.mean())
    print(f"terrain cache: rmse untrained {e_untrained:.4f} light {e_light:.4f} cached {e_cached:.4f}; "
          f"rays light {r_light} cached {r_cached}; loss {losses[0]:.4f} -> {np.mean(losses[-5:]):.4f}; "
          f"bias {bias:.4f}")
    assert np.mean(losses[-5:]) < 0.5 * losses[0]
    assert e_cached < 0.5 * e_untrained
    assert e_cached < 1.05 * e_light
    assert r_cached < r_light
    assert abs(bias) < 0.03
    cache.close()
