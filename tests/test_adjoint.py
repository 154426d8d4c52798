"""The oracle's path-replay adjoint against central finite differences of the
oracle's own renderer (common random numbers: the same Philox paths at
colour +- h).  CPU only."""
import numpy as np
import pytest

import oracle as O
from master_thesis_amd import scenes, host_lib as H
from master_thesis_amd._lib import RT_RENDER_DETERMINISTIC
from master_thesis_amd.tracer import make_settings


def _loss(sc, cam, s, w):
    img, _ = O.render(sc, cam, s)
    return float(np.sum(w.astype(np.float64) * img.astype(np.float64)))


@pytest.mark.parametrize("m,ch", [(2, 0), (17, 1), (41, 2), (50, 0), (63, 1), (64, 0), (64, 2)])
def test_oracle_adjoint_matches_central_differences(m, ch):
    sc = scenes.cornell_groups(seed=3)
    cam = H.cornell_camera(24, 24)
    s = make_settings(spp=8, max_depth=4, seed=5, flags=RT_RENDER_DETERMINISTIC)
    w = np.random.default_rng(7).uniform(-1, 1, (24, 24, 3)).astype(np.float32)
    g = O.render_adjoint(sc, cam, s, w)
    h = 1e-3
    sp, sm = scenes.cornell_groups(seed=3), scenes.cornell_groups(seed=3)
    sp.materials["color"][m, ch] += np.float32(h)
    sm.materials["color"][m, ch] -= np.float32(h)
    hp = float(sp.materials["color"][m, ch]) - float(sc.materials["color"][m, ch])
    hm = float(sc.materials["color"][m, ch]) - float(sm.materials["color"][m, ch])
    fd = (_loss(sp, cam, s, w) - _loss(sm, cam, s, w)) / (hp + hm)
    assert abs(g[m, ch] - fd) <= 2e-2 * abs(fd) + 1e-3, (g[m, ch], fd)


def test_adjoint_is_zero_where_no_path_reaches_a_light():
    w = scenes.material(scenes.WHITE)
    sc = scenes.closed_room(w, w)
    cam = H.camera((0.0, 0.0, 0.0), 0.0, -1.0, np.pi / 4, 8, 8)
    g = O.render_adjoint(sc, cam, make_settings(spp=4, max_depth=4, seed=1), np.ones((8, 8, 3), np.float32))
    assert not g.any()


def test_adjoint_of_lit_floor_closed_form():
    """Floor albedo rho under light walls E: I = 2 pi/N sum rho c1 E c2, so
    dI/drho = I / rho and dI/dE = I / E exactly (per pixel, per sample)."""
    E, rho = 2.0, 0.6
    sc = scenes.closed_room(scenes.material((rho, rho, rho)), scenes.material((E, E, E), light=True))
    cam = H.camera((0.0, 0.0, 0.0), 0.0, -1.0, np.pi / 4, 16, 16)
    s = make_settings(spp=16, max_depth=3, seed=4, flags=RT_RENDER_DETERMINISTIC)
    img, _ = O.render(sc, cam, s)
    g = O.render_adjoint(sc, cam, s, np.ones((16, 16, 3), np.float32))
    total = img.astype(np.float64).sum(axis=(0, 1))
    assert np.allclose(g[0], total / rho, rtol=1e-4)
    assert np.allclose(g[1], total / E, rtol=1e-4)
