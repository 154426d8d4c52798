"""CPU checks of the radiance-cache reference (oracle/nrc.py): the BF16
rounding against torch's bfloat16 conversion, the encoding against its closed
form, the hand-written backward pass against torch autograd and central
finite differences, and the parameter layout; the oracle's constant-cache
render (the checker of rt_render_nrc) against closed forms and the plain
render."""
import math

import numpy as np
import pytest
import torch

import oracle as O
from oracle import nrc as N
from master_thesis_amd import host_lib as H
from master_thesis_amd import scenes
from master_thesis_amd.tracer import make_settings


def _samples(n, seed=0):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 9), np.float32)
    f[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    nr = rng.normal(size=(n, 3))
    f[:, 3:6] = nr / np.linalg.norm(nr, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    f[:, 6:9] = d / np.linalg.norm(d, axis=1, keepdims=True)
    t = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    return f, t


LO, INV = np.array([-1, -1, -1], np.float32), np.array([0.5, 0.5, 0.5], np.float32)


def test_parameter_layout():
    assert N.NPARAM == 23888
    assert N.OFFW[0] == 0 and N.OFFB[0] == 96 * 64 and N.OFFW[1] == 96 * 64 + 64
    assert N.OFFB[5] + 16 == N.NPARAM


def test_bf16_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, 1e-40], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(N.bf16(x).view(np.uint32), want.view(np.uint32))


def test_encoding_closed_form():
    f = np.zeros((3, 9), np.float32)
    f[:, 0] = [-1.0, 0.0, -0.5]          # x = 0, 0.5, 0.25 in the unit box
    f[:, 3:9] = np.arange(6, dtype=np.float32)
    X = N.encode(f, LO, INV)
    assert X.shape == (3, 96)
    # x = 0: sin 0 = 0, cos 0 = 1 at every frequency
    np.testing.assert_allclose(X[0, 0:24:2], 0.0, atol=1e-12)
    np.testing.assert_allclose(X[0, 1:24:2], 1.0, atol=1e-12)
    # x = 1/2: k = 0 -> sin(pi/2) = 1; k >= 1 -> sin(2^(k-1) pi) = 0, cos = 1 for k >= 2, cos(pi) = -1 for k = 1
    assert abs(X[1, 0] - 1.0) < 1e-12 and abs(X[1, 2]) < 1e-12 and abs(X[1, 3] + 1.0) < 1e-12
    # x = 1/4: k = 0 -> sin(pi/4) = cos(pi/4) = sqrt(1/2)
    assert abs(X[2, 0] - np.sqrt(0.5)) < 1e-12 and abs(X[2, 1] - np.sqrt(0.5)) < 1e-12
    np.testing.assert_array_equal(X[:, 72:78], f[:, 3:9])
    assert np.all(X[:, 78:] == 0.0)


def _torch_loss_grad(params, feats, targets):
    """Independent gradient: torch autograd of the same network in float64."""
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    h = torch.from_numpy(N.encode(feats, LO, INV))
    for l in range(6):
        W = p[N.OFFW[l]:N.OFFW[l] + N.ND[l] * N.KD[l]].view(N.ND[l], N.KD[l])
        b = p[N.OFFB[l]:N.OFFB[l] + N.ND[l]]
        h = h @ W.T + b
        if l < 5:
            h = torch.relu(h)
    y = h[:, :3]
    t = torch.from_numpy(targets.astype(np.float64))
    n = y.shape[0]
    loss = (((y - t) ** 2) / (y.detach() ** 2 + 0.01)).sum() / n
    loss.backward()
    return float(loss.detach()) / 3, p.grad.numpy()


def test_backward_matches_autograd():
    f, t = _samples(200, 1)
    p = N.init_params(5)
    p[N.OFFB[0]:N.OFFB[0] + 64] = np.random.default_rng(2).normal(0, 0.1, 64)  # non-zero biases too
    loss, g = N.loss_and_grad(p, f, t, LO, INV, emulate=False)
    loss_t, g_t = _torch_loss_grad(p, f, t)
    assert abs(loss - loss_t) <= 1e-12 * max(1.0, loss_t)
    np.testing.assert_allclose(g, g_t, rtol=1e-9, atol=1e-12)
    # padded inputs / outputs get no gradient
    W0 = g[:N.ND[0] * N.KD[0]].reshape(64, 96)
    assert np.all(W0[:, 78:] == 0.0)
    W5 = g[N.OFFW[5]:N.OFFB[5]].reshape(16, 64)
    assert np.all(W5[3:] == 0.0) and np.all(g[N.OFFB[5] + 3:] == 0.0)


def test_backward_matches_finite_differences():
    """Central differences of a float64 forward pass on 36 parameters (six per
    layer).  The relative-L2 loss treats its denominator y^2 + 0.01 as a
    constant, so the differenced loss holds it at the base point's value."""
    f, t = _samples(16, 4)
    p = N.init_params(7).astype(np.float64)
    rng = np.random.default_rng(9)
    _, g = N.loss_and_grad(p.astype(np.float32), f, t, LO, INV, emulate=False)

    # the oracle's forward casts parameters to float32; differentiate a float64 copy
    def loss64(pp, den):
        W, b = N.unpack(pp)
        h = N.encode(f, LO, INV)
        for l in range(6):
            h = h @ W[l].T + b[l]
            if l < 5:
                h = np.maximum(h, 0.0)
        y = h[:, :3]
        return float(np.sum((y - t) ** 2 / den) / f.shape[0])

    y0, _ = N.forward(p.astype(np.float32), f, LO, INV, emulate=False)
    den = y0 ** 2 + 0.01
    p32 = p.astype(np.float32).astype(np.float64)
    idx = np.concatenate([rng.integers(N.OFFW[l], N.OFFB[l] + N.ND[l], 6) for l in range(6)])
    for i in idx:
        if g[i] == 0.0:
            continue
        h = 1e-6
        e = np.zeros_like(p32)
        e[i] = h
        fd = (loss64(p32 + e, den) - loss64(p32 - e, den)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_emulated_forward_close_to_exact():
    f, _ = _samples(300, 2)
    p = N.init_params(1)
    ye, _ = N.forward(p, f, LO, INV, emulate=True)
    yx, _ = N.forward(p, f, LO, INV, emulate=False)
    err = np.abs(ye - yx)
    assert err.max() < 0.05 * max(1.0, np.abs(yx).max())
    assert err.max() > 0.0  # the emulation rounds somewhere


# ---- the constant-cache render of the oracle (the checker of rt_render_nrc)


def _down_camera(w=32, h=32):
    # centre of the closed room, looking straight down at the floor
    return H.camera((0.0, 0.0, 0.0), 0.0, -1.0, math.pi / 4, w, h)


def test_cached_render_is_pi_rho_c():
    """Camera rays hit the floor (albedo rho); every bounce from it hits a
    non-emitting wall, the first-bounce vertex, where the sample adds
    throughput rho * cos * c: pixel mean pi * rho * c (cos ~ U[0, 1])."""
    rho = 0.6
    sc = scenes.closed_room(scenes.material((rho, rho, rho)), scenes.material(scenes.WHITE))
    cam = _down_camera()
    c = np.array([0.3, 0.5, 0.7], np.float32)
    N = 16
    rgb, nq, nr = O.render_cached(sc, cam, make_settings(spp=N, max_depth=4, seed=5), c, q_vertex=1, train_every=0)
    P = cam.width * cam.height
    assert nq == N * P and nr == 0
    for ch in range(3):
        sd = 2 * math.pi * rho * c[ch] / math.sqrt(12 * N * P)
        assert abs(rgb[..., ch].mean() - math.pi * rho * c[ch]) < 4 * sd
    # a zero cache leaves nothing (the room has no light)
    rgb0, _, _ = O.render_cached(sc, cam, make_settings(spp=N, max_depth=4, seed=5), np.zeros(3), 1, 0)
    assert not rgb0.any()
    # negative predictions count as zero
    rgbn, _, _ = O.render_cached(sc, cam, make_settings(spp=N, max_depth=4, seed=5), -c, 1, 0)
    assert not rgbn.any()


def test_cached_render_skips_lights_and_last_depth():
    """A first-bounce vertex on a light ends the path as in the plain render
    (no query); so does a vertex at the last depth."""
    rho, E = 0.6, 2.0
    sc = scenes.closed_room(scenes.material((rho, rho, rho)), scenes.material((E, E, E), light=True))
    cam = _down_camera(16, 16)
    s = make_settings(spp=8, max_depth=4, seed=6)
    plain, _ = O.render(sc, cam, s)
    cached, nq, nr = O.render_cached(sc, cam, s, np.array([9.0, 9.0, 9.0]), 1, 0)
    assert nq == 0 and nr == 0
    np.testing.assert_array_equal(cached[0], plain)
    sc2 = scenes.closed_room(scenes.material((rho, rho, rho)), scenes.material(scenes.WHITE))
    s2 = make_settings(spp=8, max_depth=2, seed=6)
    cached2, nq2, _ = O.render_cached(sc2, cam, s2, np.array([9.0, 9.0, 9.0]), 1, 0)
    assert nq2 == 0 and not cached2.any()


def test_cached_render_training_samples_trace_on():
    """Training samples follow the plain render's paths; with train_every = 1
    every sample is one and the image is the plain image bit for bit."""
    sc = scenes.cornell_box()
    cam = H.cornell_camera(24, 24)
    s = make_settings(spp=4, max_depth=5, seed=7)
    plain, _ = O.render(sc, cam, s)
    cached, nq, nr = O.render_cached(sc, cam, s, np.array([1.0, 1.0, 1.0]), 1, 1)
    assert nq == 0 and nr > 0
    np.testing.assert_array_equal(cached[0], plain)
    # train_every = 4: about a quarter of the samples that reach the vertex train
    _, nq4, nr4 = O.render_cached(sc, cam, s, np.array([1.0, 1.0, 1.0]), 1, 4)
    assert nq4 + nr4 == nr
    assert abs(nr4 / nr - 0.25) < 0.05
    # query off: every sample traces its whole path
    off, nq0, _ = O.render_cached(sc, cam, s, np.array([1.0, 1.0, 1.0]), 1, 4, query=False)
    assert nq0 == 0
    np.testing.assert_array_equal(off[0], plain)


# ---- light sampling of the cache mode (oracle)

def _sun_settings(spp, theta_deg=30.0, alpha_deg=5.0, S=40.0, seed=3, max_depth=3):
    t = math.radians(theta_deg)
    return make_settings(spp=spp, max_depth=max_depth, seed=seed, flags=0x4,
                         sun_dir=(math.sin(t), math.cos(t), 0.0), sun_radius_deg=alpha_deg,
                         sun_radiance=(S, S, S))


def test_sun_sampling_matches_closed_form():
    """Open ground (albedo rho) under a sun of half-angle alpha at zenith
    angle theta: pixel = rho * S * pi * sin^2(alpha) * cos(theta) for the
    reference estimator; light sampling has the same mean at far lower
    variance."""
    rho, S, th, al = 0.5, 40.0, 30.0, 5.0
    sc = scenes.ground_plane((rho, rho, rho))
    cam = H.camera((0.0, 1.0, 0.0), 0.0, -1.0, math.pi / 4, 32, 32)
    want = rho * S * math.pi * math.sin(math.radians(al)) ** 2 * math.cos(math.radians(th))
    nee, nq, nr = O.render_cached(sc, cam, _sun_settings(4), np.zeros(3), q_vertex=1, train_every=0,
                                  query=False, nee=True)
    assert nq == 0 and nr == 0
    assert abs(nee.mean() - want) < 2e-3 * want
    plain, _ = O.render(sc, cam, _sun_settings(256))
    # plain per-sample value 2 pi rho S cos, non-zero with probability 1 - cos(alpha) (the cone's share of
    # the hemisphere): its standard deviation is below 2 pi rho S sqrt(1 - cos(alpha))
    p_hit = 1.0 - math.cos(math.radians(al))
    sd = 2 * math.pi * rho * S * math.sqrt(p_hit) / math.sqrt(256 * 32 * 32)
    assert abs(plain.mean() - want) < 4 * sd
    assert nee.std() < 0.2 * plain.std()


def test_light_sampling_lit_floor_is_half_pi_rho_E():
    """Closed room, floor albedo rho, emissive walls E: the reference
    estimator's mean pi rho E / 2 (cosines at the floor and at the light),
    reproduced by sampling the walls' area."""
    rho, E = 0.6, 2.0
    sc = scenes.closed_room(scenes.material((rho, rho, rho)), scenes.material((E, E, E), light=True))
    cam = _down_camera()
    N = 16
    nee, _, _ = O.render_cached(sc, cam, make_settings(spp=N, max_depth=4, seed=4), np.zeros(3), 1, 0,
                                query=False, nee=True)
    want = math.pi * rho * E / 2
    P = cam.width * cam.height
    # the estimate per sample is bounded by a few times its mean; a 1% band is > 6 sigma here
    assert abs(nee.mean() - want) < 0.01 * want
    plain, _ = O.render(sc, cam, make_settings(spp=N, max_depth=4, seed=4))
    sd = 2 * math.pi * rho * E * math.sqrt(1 / 9 - 1 / 16) / math.sqrt(N * P)
    assert abs(plain.mean() - want) < 4 * sd


def test_light_sampling_cornell_matches_plain_mean():
    """Cornell box, whole frame: the light-sampled image converges to the
    plain image (same expectation), checked on the frame mean."""
    sc = scenes.cornell_box()
    cam = H.cornell_camera(24, 24)
    plain, _ = O.render(sc, cam, make_settings(spp=1024, max_depth=5, seed=1))
    nee, _, _ = O.render_cached(sc, cam, make_settings(spp=256, max_depth=5, seed=2), np.zeros(3), 1, 0,
                                query=False, nee=True)
    assert abs(nee.mean() - plain.mean()) < 0.03 * plain.mean()
    # and it is the lower-variance estimate: its per-pixel error against the plain 1024-spp image is smaller
    plain2, _ = O.render(sc, cam, make_settings(spp=256, max_depth=5, seed=3))
    assert np.sqrt(np.mean((nee[0] - plain) ** 2)) < np.sqrt(np.mean((plain2 - plain) ** 2))


def test_training_records_hold_the_rest_of_the_path():
    """A record's target is the rest of its path's contribution from the
    query vertex on.  With the query vertex at the primary hit (index 0, an
    oracle-only setting) and every sample a training sample, the targets
    summed per pixel times 2 pi / N reproduce the plain image; no record is
    made at a vertex that is a light."""
    rho, E = 0.6, 2.0
    sc = scenes.closed_room(scenes.material((rho, rho, rho)), scenes.material((E, E, E), light=True))
    cam = _down_camera(8, 8)
    s = make_settings(spp=4, max_depth=4, seed=4)
    # vertex 1 is always a light here: no records
    _, _, nr, f, t, _ = O.render_cached(sc, cam, s, np.zeros(3), 1, 1, records=True)
    assert nr == 0 and f.shape[0] == 0
    # q = 0: every sample records at the floor, its target is its whole contribution
    rgb, _, nr0, f0, t0, _ = O.render_cached(sc, cam, s, np.zeros(3), 0, 1, records=True)
    assert nr0 == 8 * 8 * 4
    plain, _ = O.render(sc, cam, s)
    per_pixel = t0.reshape(8 * 8, 4, 3).sum(axis=1).reshape(8, 8, 3)
    scale = np.float32(2 * math.pi * (1.0 / 4))
    np.testing.assert_allclose(per_pixel * scale, plain, rtol=1e-5)
    np.testing.assert_array_equal(rgb[0], plain)


@pytest.mark.parametrize("nee", [False, True])
def test_record_target_is_the_rest_of_the_path(nee):
    """One sample per pixel, query vertex 1.  The random numbers do not
    depend on whether a sample queries or trains, so per pixel, with I0 the
    image with a zero cache, I1 with a cache of 1 and It with every sample
    training: I1 - I0 = s * T1 (s = 2 pi, T1 the throughput up to the query
    vertex) and It - I0 = s * T1 * target.  The records must satisfy that."""
    sc = scenes.cornell_box()
    cam = H.cornell_camera(24, 24)
    st = make_settings(spp=1, max_depth=6, seed=21)
    I0, _, _ = O.render_cached(sc, cam, st, np.zeros(3), 1, 0, nee=nee)
    I1, nq, _ = O.render_cached(sc, cam, st, np.ones(3), 1, 0, nee=nee)
    It, _, nr, f, t, keys = O.render_cached(sc, cam, st, np.ones(3), 1, 1, nee=nee, records=True)
    assert nr == nq > 100
    w = (I1[0] - I0[0]).reshape(-1, 3)[keys].astype(np.float64)
    dt = (It[0] - I0[0]).reshape(-1, 3)[keys].astype(np.float64)
    # pixels that did not reach the query vertex have no record and identical images
    rest = np.setdiff1d(np.arange(24 * 24), keys)
    np.testing.assert_array_equal(I1[0].reshape(-1, 3)[rest], I0[0].reshape(-1, 3)[rest])
    np.testing.assert_array_equal(It[0].reshape(-1, 3)[rest], I0[0].reshape(-1, 3)[rest])
    scale = 1e-6 * (1.0 + np.abs(dt) + np.abs(I0[0].reshape(-1, 3)[keys]))
    assert np.all(np.abs(w * t - dt) <= scale + 1e-5 * np.abs(w * t)), np.max(np.abs(w * t - dt))
    assert (t > 0).any()
