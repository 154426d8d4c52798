"""GPU post-filters (rt_denoise_*) bit-identical to the oracle's on random
images, host and device pointers; the SoftwareRenderer-compatible command
line chains GaussDenoiser and MedianDenoiser entries after the renderer."""
import subprocess

import numpy as np
import pytest

import oracle as O
from master_thesis_amd import host_lib as H
from master_thesis_amd import images as I
from master_thesis_amd._lib import RTError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,sigma", [(1, 1.0), (3, 0.8), (5, 1.5), (19, 3.0)])
def test_gauss_matches_oracle(tracer, k, sigma):
    rng = np.random.default_rng(k)
    img = rng.uniform(0, 3, (37, 53, 3)).astype(np.float32)
    got = tracer.denoise_gauss(img, k, sigma)
    want = O.denoise_gauss(img, k, sigma)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("k,sigma", [(5, 1.5), (19, 3.0)])
def test_filters_match_oracle_with_interior_tiles(tracer, k, sigma):
    """An image large enough for tiles away from the border (which skip the
    per-tap bounds test) and ragged at both edges: both filters
    bit-identical to the oracle's."""
    rng = np.random.default_rng(100 + k)
    img = rng.uniform(-1, 3, (157, 301, 3)).astype(np.float32)
    got = tracer.denoise_gauss(img, k, sigma)
    want = O.denoise_gauss(img, k, sigma)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    got = tracer.denoise_median(img)
    want = O.denoise_median(img)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_median_matches_oracle_on_device(tracer):
    import torch
    rng = np.random.default_rng(5)
    img = rng.uniform(-1, 3, (70, 45, 3)).astype(np.float32)
    got = tracer.denoise_median(torch.from_numpy(img).cuda())
    torch.cuda.synchronize()
    want = O.denoise_median(img)
    assert np.array_equal(got.cpu().numpy().view(np.uint32), want.view(np.uint32))
    got = tracer.denoise_median(img)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_filter_argument_errors(tracer):
    img = np.zeros((8, 8, 3), np.float32)
    for k in (0, 4, 21):
        with pytest.raises(RTError):
            tracer.denoise_gauss(img, k, 1.0)
    with pytest.raises(RTError):
        tracer.denoise_gauss(img, 3, 0.0)


def test_cli_chains_denoisers(tracer, tmp_path):
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
        "\tcolor(15, 15, 15)\n")
    (tmp_path / "r.renderer").write_text("PathTracer\n\tvisionJumpCount(5)\n\traysPerPixel(8)\n\tseed(3)\n"
                                         "GaussDenoiser\n\tkernelSize(5)\n\tsigma(1.5)\n"
                                         "MedianDenoiser\n")
    (tmp_path / "c.camera").write_text("Camera\n\tposition(3.4, 0, 0)\n\tangle(1, 0)\n")
    out = tmp_path / "o.ppm"
    p = subprocess.run([str(B.CLI), str(tmp_path / "r.renderer"), str(tmp_path / "scene" / "box.scene"), "64", "48",
                        "2", str(tmp_path / "c.camera"), str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    img = I.read_ppm(out)
    sc = H.load_scene(str(tmp_path / "scene" / "box.scene"), str(tmp_path / "obj"))
    tracer.upload(sc)
    tracer.build()
    cam = H.camera_from_file(str(tmp_path / "c.camera"), 64, 48)
    _, s = H.renderer_from_file(str(tmp_path / "r.renderer"))
    rgb, _ = tracer.render(cam, s)
    want = I.to_rgb8(tracer.denoise_median(tracer.denoise_gauss(rgb, 5, 1.5)))
    assert np.array_equal(img, want)
    (tmp_path / "bad.renderer").write_text("PathTracer\n\tvisionJumpCount(5)\n\traysPerPixel(1)\nBilateral\n")
    p = subprocess.run([str(B.CLI), str(tmp_path / "bad.renderer"), str(tmp_path / "scene" / "box.scene"), "8", "8",
                        "1", str(tmp_path / "c.camera"), str(tmp_path / "x.ppm")], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 1 and "getDenoiser not found" in p.stderr
