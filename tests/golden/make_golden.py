"""Writes the golden vectors of tests/golden/ from the reference compiled from
its own sources (oracle/_ref/libref.so; needs the reference present, so it
runs in the build container only).  The vectors are data: rays, the
reference's answers, and camera matrices.

    python tests/golden/make_golden.py            (all vectors)
    python tests/golden/make_golden.py occlusion  (the occlusion vectors only)
"""
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import ref as R  # noqa: E402
from master_thesis_amd import scenes  # noqa: E402
from master_thesis_amd._lib import RAY_DTYPE  # noqa: E402

CAMERAS = [((3.4, 0, 0), 1, 0), ((4, 0, 0), 1, 0), ((8.29397, 0, -1.6169), 0.783549, -0.0445634),
           ((0, 60, 0), 1, -1), ((1, 2, 3), 1.7, 0.3), ((-2, 0.5, 1), -1.3, 1.4)]


def main():
    assert R.available(), "build oracle/_ref first (make -C oracle ref)"
    sc = scenes.cornell_box()
    rng = np.random.default_rng(20231203)
    n = 4096
    rays = np.zeros(n, RAY_DTYPE)
    rays["o"] = rng.uniform(-2.7, 2.7, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.inf
    hit, p, nr, obj = R.RefScene(sc).trace(rays)
    np.savez_compressed(HERE / "ref_cornell_trace.npz", rays=rays.view(np.float32).reshape(n, 8), hit=hit,
                        pos=p, normal=nr, obj=obj)
    mats = []
    for pos, a0, a1 in CAMERAS:
        vi, pi, og = R.camera(pos, a0, a1, 320, 200)
        mats.append(np.concatenate([np.asarray(pos, np.float32), [a0, a1], vi, pi, og]).astype(np.float32))
    np.savez_compressed(HERE / "ref_cameras.npz", rows=np.stack(mats), width=320, height=200,
                        fov=np.float32(math.pi / 4))
    print("wrote", HERE / "ref_cornell_trace.npz", HERE / "ref_cameras.npz")
    occlusion()


def occlusion():
    """Scene::isOccluded on the Cornell box: random segments, segments ending
    exactly at the reference's own hit points (the inclusive bound), and
    segments stopping short of them."""
    sc = scenes.cornell_box()
    ref = R.RefScene(sc)
    rng = np.random.default_rng(20231204)
    n = 2048
    segs = [np.concatenate([rng.uniform(-2.7, 2.7, (n, 3)), rng.uniform(-2.7, 2.7, (n, 3))], axis=1)]
    rays = np.zeros(n, RAY_DTYPE)
    rays["o"] = rng.uniform(-2.7, 2.7, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.inf
    hit, p, _, _ = ref.trace(rays)
    o = rays["o"][hit]
    segs.append(np.concatenate([o, p[hit]], axis=1))
    segs.append(np.concatenate([o, o + np.float32(0.999) * (p[hit] - o)], axis=1))
    segs = np.concatenate(segs).astype(np.float32)
    occ = ref.occluded(segs)
    np.savez_compressed(HERE / "ref_cornell_occluded.npz", segs=segs, occluded=occ)
    print("wrote", HERE / "ref_cornell_occluded.npz", segs.shape[0], "segments,", int(occ.sum()), "occluded")


if __name__ == "__main__":
    if sys.argv[1:] == ["occlusion"]:
        occlusion()
    else:
        main()
