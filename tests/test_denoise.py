"""The oracle's post-filters (the Vulkan renderer's GaussDenoiser and
MedianDenoiser) against closed forms: the weights' formula (centred on the
last index, as the reference computes them), the impulse response, skipped
taps at the border, and the 3x3 median."""
import math

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("k,sigma", [(1, 1.0), (3, 0.7), (5, 1.5), (19, 4.0)])
def test_gauss_weights_closed_form(k, sigma):
    g = O.gauss_weights(k, sigma)
    raw = np.array([math.exp(-((i - (k - 1)) ** 2) / (2 * sigma * sigma)) for i in range(k)])
    np.testing.assert_allclose(g, raw / raw.sum(), rtol=2e-6)
    assert abs(float(g.astype(np.float64).sum()) - 1.0) < 1e-5
    assert int(np.argmax(g)) == k - 1  # the reference centres the kernel on its last tap


def test_gauss_impulse_response_and_constant_interior():
    k, s = 5, 1.2
    g = O.gauss_weights(k, s)
    img = np.zeros((15, 17, 3), np.float32)
    img[7, 8] = [1.0, 2.0, 4.0]
    out = O.denoise_gauss(img, k, s)
    # out(p) = sum in(p + (x, y)) g[x + 2] g[y + 2]: the impulse lands at p = q - (x, y)
    for x in range(-2, 3):
        for y in range(-2, 3):
            np.testing.assert_array_equal(out[7 - y, 8 - x], np.float32([1, 2, 4]) * g[x + 2] * g[y + 2])
    assert np.count_nonzero(out[..., 0]) == 25
    c = np.full((15, 17, 3), 0.37, np.float32)
    out = O.denoise_gauss(c, k, s)
    np.testing.assert_allclose(out[2:-2, 2:-2], 0.37, rtol=2e-6)
    # the corner keeps only the taps inside the image (no renormalisation)
    inside = sum(g[x + 2] * g[y + 2] for x in range(0, 3) for y in range(0, 3))
    np.testing.assert_allclose(out[0, 0], 0.37 * inside, rtol=2e-6)


def test_median_is_the_3x3_median():
    rng = np.random.default_rng(3)
    img = rng.uniform(-1, 2, (13, 11, 3)).astype(np.float32)
    out = O.denoise_median(img)
    for y in range(1, 12):
        for x in range(1, 10):
            for c in range(3):
                assert out[y, x, c] == np.median(img[y - 1:y + 2, x - 1:x + 2, c])
    # border: the missing taps are 0
    pad = np.zeros((15, 13, 3), np.float32)
    pad[1:-1, 1:-1] = img
    for (y, x) in ((0, 0), (0, 5), (12, 10), (6, 0)):
        for c in range(3):
            assert out[y, x, c] == np.median(pad[y:y + 3, x:x + 3, c])
