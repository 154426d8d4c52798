"""GPU adjoint: agrees with the oracle's CPU adjoint (same paths; only the
accumulation order of the per-material sums differs), with central finite
differences of the GPU renderer, and through torch.autograd."""
import numpy as np
import pytest

import oracle as O
from master_thesis_amd import scenes, host_lib as H
from master_thesis_amd._lib import RT_RENDER_DETERMINISTIC
from master_thesis_amd.tracer import make_settings

pytestmark = pytest.mark.gpu


def _setup(tracer, sc):
    tracer.upload(sc)
    tracer.build(4)


def test_gpu_adjoint_matches_oracle(tracer):
    sc = scenes.cornell_groups(seed=11)
    _setup(tracer, sc)
    cam = H.cornell_camera(40, 40)
    s = make_settings(spp=16, max_depth=4, seed=13, flags=RT_RENDER_DETERMINISTIC)
    w = np.random.default_rng(1).uniform(-1, 1, (40, 40, 3)).astype(np.float32)
    g = tracer.render_adjoint(cam, s, w)
    want = O.render_adjoint(sc, cam, s, w)
    scale = np.abs(want).max()
    assert np.allclose(g, want, rtol=1e-4, atol=1e-5 * scale), np.abs(g - want).max()


def test_gpu_adjoint_with_culling_steps_matches_oracle(tracer):
    """Camera outside a closed room: the primary rays pass walls from behind
    before their first vertex, which later samples take from the item's
    cache; the gradient still equals the oracle's."""
    import math
    sc = scenes.closed_room(scenes.material((0.6, 0.5, 0.4)), scenes.material((1.5, 1.2, 1.0), light=True))
    _setup(tracer, sc)
    cam = H.camera((3.0, 0.3, 0.2), 1.0, 0.05, math.pi / 2, 24, 20)
    s = make_settings(spp=8, max_depth=4, seed=17, flags=RT_RENDER_DETERMINISTIC)
    w = np.random.default_rng(2).uniform(-1, 1, (20, 24, 3)).astype(np.float32)
    g = tracer.render_adjoint(cam, s, w)
    want = O.render_adjoint(sc, cam, s, w)
    scale = np.abs(want).max()
    assert scale > 0
    assert np.allclose(g, want, rtol=1e-4, atol=1e-5 * scale), np.abs(g - want).max()


def test_gpu_adjoint_throughput_mode_and_many_materials(tracer):
    """Throughput mode (a pixel's samples split over several slots) gives the
    oracle's gradient; with more than 1024 materials the per-workgroup LDS
    sums do not fit and the gradient goes to global atomics directly: the
    same values, and 0 for materials no triangle uses."""
    import dataclasses
    sc = scenes.cornell_groups(seed=11)
    cam = H.cornell_camera(40, 40)
    s = make_settings(spp=16, max_depth=4, seed=13)
    w = np.random.default_rng(1).uniform(-1, 1, (40, 40, 3)).astype(np.float32)
    want = O.render_adjoint(sc, cam, s, w)
    scale = np.abs(want).max()
    _setup(tracer, sc)
    g_lds = tracer.render_adjoint(cam, s, w)
    assert np.allclose(g_lds, want, rtol=1e-4, atol=1e-5 * scale), np.abs(g_lds - want).max()
    M = sc.materials.shape[0]
    extra = np.repeat(sc.materials[:1], 1100 - M, axis=0)
    big = dataclasses.replace(sc, materials=np.concatenate([sc.materials, extra]))
    _setup(tracer, big)
    g_atom = tracer.render_adjoint(cam, s, w)
    assert g_atom.shape == (1100, 3)
    assert np.allclose(g_atom[:M], want, rtol=1e-4, atol=1e-5 * scale), np.abs(g_atom[:M] - want).max()
    assert np.all(g_atom[M:] == 0.0)


@pytest.mark.parametrize("m,ch", [(3, 0), (22, 1), (60, 2), (64, 1)])
def test_gpu_adjoint_matches_central_differences(tracer, m, ch):
    sc = scenes.cornell_groups(seed=2)
    _setup(tracer, sc)
    cam = H.cornell_camera(32, 32)
    s = make_settings(spp=16, max_depth=4, seed=21, flags=RT_RENDER_DETERMINISTIC)
    w = np.random.default_rng(9).uniform(-1, 1, (32, 32, 3)).astype(np.float32)
    g = tracer.render_adjoint(cam, s, w)
    base = sc.materials["color"].copy()
    h = np.float32(1e-3)

    def loss(c):
        tracer.set_material_colors(c)
        img, _ = tracer.render(cam, s)
        return float(np.sum(w.astype(np.float64) * img.astype(np.float64)))

    cp, cm = base.copy(), base.copy()
    cp[m, ch] += h
    cm[m, ch] -= h
    fd = (loss(cp) - loss(cm)) / (float(cp[m, ch]) - float(cm[m, ch]))
    tracer.set_material_colors(base)
    assert abs(g[m, ch] - fd) <= 2e-2 * abs(fd) + 1e-3, (g[m, ch], fd)


def test_full_configuration_l2_gradient(tracer):
    """The differentiable workload at its full size: 64 albedo groups,
    256 x 256, 32 spp, depth 4, L2 loss against a render with seed-drawn
    true albedos.  The GPU's gradient (forward render, then the adjoint of
    w = 2 (image - target)) equals the oracle's adjoint of the same w, and
    central differences of the GPU renderer's loss (h = 1e-3, common random
    numbers) for the group with the largest gradient and one near the
    median."""
    cam = H.cornell_camera(256, 256)
    s = make_settings(spp=32, max_depth=4, seed=13, flags=RT_RENDER_DETERMINISTIC)
    _setup(tracer, scenes.cornell_groups(seed=99))
    target, _ = tracer.render(cam, s)
    target = target.astype(np.float64)
    sc = scenes.cornell_groups(seed=11)
    _setup(tracer, sc)
    img, _ = tracer.render(cam, s)
    w = (2.0 * (img.astype(np.float64) - target)).astype(np.float32)
    g = tracer.render_adjoint(cam, s, w)
    want = O.render_adjoint(sc, cam, s, w)
    scale = np.abs(want).max()
    assert scale > 0
    assert np.allclose(g, want, rtol=1e-4, atol=1e-5 * scale), np.abs(g - want).max()
    base = sc.materials["color"].copy()
    h = np.float32(1e-3)

    def loss(c):
        tracer.set_material_colors(c)
        im, _ = tracer.render(cam, s)
        return float(np.sum((im.astype(np.float64) - target) ** 2))

    flat = np.abs(want[:64]).ravel()
    picks = [int(np.argmax(flat)), int(np.argsort(flat)[flat.size // 2])]
    for k in picks:
        m, ch = divmod(k, 3)
        cp, cm = base.copy(), base.copy()
        cp[m, ch] += h
        cm[m, ch] -= h
        fd = (loss(cp) - loss(cm)) / (float(cp[m, ch]) - float(cm[m, ch]))
        assert abs(g[m, ch] - fd) <= 2e-2 * abs(fd) + 1e-4 * scale, (m, ch, g[m, ch], fd)
    tracer.set_material_colors(base)


def test_autograd_function_gradient_and_descent(tracer):
    import torch
    from master_thesis_amd import diff

    true = scenes.cornell_groups(seed=31)
    _setup(tracer, true)
    cam = H.cornell_camera(32, 32)
    s = make_settings(spp=16, max_depth=4, seed=41)
    colors_true = torch.from_numpy(true.materials["color"].copy()).cuda()
    with torch.no_grad():
        target = diff.render(colors_true, tracer, cam, s).clone()
    init = scenes.cornell_groups(seed=32).materials["color"].copy()
    init[64] = true.materials["color"][64]  # the light stays fixed
    colors = torch.from_numpy(init).cuda().requires_grad_(True)
    img = diff.render(colors, tracer, cam, s)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    # the same gradient through the C ABI directly
    dimg = (2 * (img.detach() - target)).contiguous()
    direct = tracer.render_adjoint(cam, s, dimg)
    torch.cuda.synchronize()
    assert torch.allclose(colors.grad, direct, rtol=1e-5, atol=1e-6)
    # a few descent steps on the albedos reduce the loss
    opt = torch.optim.Adam([colors], lr=2e-2)
    first = float(loss.detach())
    for _ in range(15):
        opt.zero_grad()
        l2 = ((diff.render(colors, tracer, cam, s) - target) ** 2).sum()
        l2.backward()
        colors.grad[64] = 0.0
        opt.step()
        with torch.no_grad():
            colors[:64].clamp_(0.0, 1.0)  # albedos only; the light's colour is its emission
    assert float(l2.detach()) < 0.5 * first, (float(l2.detach()), first)
