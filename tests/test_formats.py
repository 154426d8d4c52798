"""Scene formats and image files at scale: binary STL (y/z swap, normal
smoothing within pi/3, the reference's corner-grouping order) against closed
forms and the compiled reference's STLLoader; scene files naming STL meshes;
PPM / PNG files; master.py's image-difference metric."""
import math
import struct

import numpy as np
import pytest

from master_thesis_amd import host_lib as H
from master_thesis_amd import images as I
from master_thesis_amd import scenes
from master_thesis_amd._lib import RTError


def write_stl(path, verts, normals, header=b"binary stl"):
    """verts (T, 3, 3) and facet normals (T, 3) in the file's own axes."""
    verts = np.asarray(verts, np.float32)
    normals = np.asarray(normals, np.float32)
    out = bytearray(header.ljust(80, b" ")[:80])
    out += struct.pack("<I", verts.shape[0])
    for t in range(verts.shape[0]):
        out += normals[t].tobytes() + verts[t].reshape(9).tobytes() + b"\0\0"
    path.write_bytes(bytes(out))


def corner_data(scene):
    """Per-triangle corner positions and normals, (T, 3, 3) each."""
    return scene.pos[scene.idx], scene.nrm[scene.idx]


def file_axes(v):
    """Our (x, y, z) as the STL file stores it (x, z, y): the swap is its own inverse."""
    v = np.asarray(v, np.float32)
    return v[..., [0, 2, 1]]


def cube_mesh():
    """Unit cube, 12 outward triangles with flat facet normals (our axes)."""
    tris, nrms = [], []
    for axis in range(3):
        for s in (0.0, 1.0):
            n = np.zeros(3)
            n[axis] = 1.0 if s else -1.0
            u, v = [a for a in range(3) if a != axis]
            q = []
            for (a, b) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                p = np.zeros(3)
                p[axis], p[u], p[v] = s, a, b
                q.append(p)
            for tri in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
                tris.append(tri)
                nrms.append(n)
    return np.array(tris, np.float32), np.array(nrms, np.float32)


def sphere_mesh(level=2):
    v, f = scenes.icosphere(level)
    tris = v[f].astype(np.float32)
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return tris, n.astype(np.float32)


def fan_mesh():
    """Three facets meeting at the origin with normals at 0, 50 and 100
    degrees from +y (in the xy plane), in that triangle order."""
    tris, nrms = [], []
    for k, deg in enumerate((0.0, 50.0, 100.0)):
        a = math.radians(deg)
        n = np.array([math.sin(a), math.cos(a), 0.0])
        t1 = np.array([math.cos(a), -math.sin(a), 0.0])
        t2 = np.array([0.0, 0.0, 1.0]) * (1 + k)
        tris.append([np.zeros(3), t1, t2])
        nrms.append(n)
    return np.array(tris, np.float32), np.array(nrms, np.float32)


def test_stl_axes_header_and_count(tmp_path):
    p = tmp_path / "t.stl"
    write_stl(p, [[[1, 2, 3], [4, 5, 6], [7, 8, 9]]], [[0, 0, 1]], header=b"solid lookalike header")
    s = H.load_mesh(str(p))
    pos, nrm = corner_data(s)
    np.testing.assert_array_equal(pos[0], [[1, 3, 2], [4, 6, 5], [7, 9, 8]])
    np.testing.assert_array_equal(nrm[0], [[0, 1, 0]] * 3)
    p.write_bytes(p.read_bytes()[:100])
    with pytest.raises(RTError, match="truncated"):
        H.load_mesh(str(p))
    with pytest.raises(RTError, match="Could not load STL"):
        H.load_mesh(str(tmp_path / "missing.stl"))


def test_stl_cube_keeps_hard_edges(tmp_path):
    """Faces meet at 90 degrees (> pi/3): each of the 8 corners splits into
    3 vertices, one per face, each with its face's normal."""
    tris, nrms = cube_mesh()
    write_stl(tmp_path / "c.stl", file_axes(tris), file_axes(nrms))
    s = H.load_mesh(str(tmp_path / "c.stl"))
    assert s.pos.shape == (24, 3) and s.idx.shape == (12, 3)
    pos, nrm = corner_data(s)
    np.testing.assert_array_equal(pos, tris)
    np.testing.assert_array_equal(nrm, np.repeat(nrms[:, None, :], 3, axis=1))


def test_stl_sphere_is_smoothed(tmp_path):
    """Neighbouring facets of an icosphere differ by < 60 degrees: one vertex
    per position, its normal the mean of the facet normals around it."""
    tris, nrms = sphere_mesh(2)
    write_stl(tmp_path / "s.stl", file_axes(tris), file_axes(nrms))
    s = H.load_mesh(str(tmp_path / "s.stl"))
    v, _ = scenes.icosphere(2)
    assert s.pos.shape[0] == v.shape[0]
    pos, nrm = corner_data(s)
    np.testing.assert_array_equal(pos, tris)
    for p in np.unique(tris.reshape(-1, 3), axis=0)[:40]:
        touching = np.any(np.all(tris == p, axis=2), axis=1)
        want = nrms[touching].astype(np.float64).mean(axis=0)
        got = nrm[np.all(pos == p, axis=2)]
        np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), atol=2e-7)


def test_stl_grouping_starts_from_the_last_corner(tmp_path):
    """Normals at 0, 50, 100 degrees: the last corner (100) opens the first
    group, 0 is too far from it and opens a second, 50 joins the first; so
    the 50-degree facet's corner gets the mean of 50 and 100 (a first-corner
    start would pair 0 with 50 instead)."""
    tris, nrms = fan_mesh()
    write_stl(tmp_path / "f.stl", file_axes(tris), file_axes(nrms))
    s = H.load_mesh(str(tmp_path / "f.stl"))
    _, nrm = corner_data(s)
    np.testing.assert_allclose(nrm[1, 0], (nrms[1] + nrms[2]) / 2, atol=1e-7)
    np.testing.assert_allclose(nrm[2, 0], (nrms[1] + nrms[2]) / 2, atol=1e-7)
    np.testing.assert_array_equal(nrm[0, 0], nrms[0])


def _ref_available():
    import oracle.ref as R
    return R.available()


@pytest.mark.skipif(not _ref_available(), reason="oracle/_ref/libref.so is not built (needs /root/reference)")
def test_stl_matches_compiled_reference(tmp_path):
    """Every triangle corner's position and normal bit-identical to the
    compiled reference's STLLoader (its vertex numbering follows its hash
    map and is not compared)."""
    import oracle.ref as R
    (tmp_path / "res" / "stl").mkdir(parents=True)
    (tmp_path / "work").mkdir()
    rng = np.random.default_rng(4)
    # a bumpy 12 x 12 patch: facet normals from the geometry plus noise, so
    # some neighbours fall on either side of the threshold
    n = 12
    hgt = rng.normal(0, 0.35, (n, n))
    pts = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="xy"), -1).astype(np.float64)
    P = np.concatenate([pts[..., :1], hgt[..., None], pts[..., 1:]], -1)
    patch = []
    for i in range(n - 1):
        for j in range(n - 1):
            patch.append([P[i, j], P[i + 1, j], P[i, j + 1]])
            patch.append([P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]])
    patch = np.array(patch, np.float32)
    pn = np.cross(patch[:, 1] - patch[:, 0], patch[:, 2] - patch[:, 0])
    pn = (pn / np.linalg.norm(pn, axis=1, keepdims=True) + rng.normal(0, 0.05, pn.shape)).astype(np.float32)
    for name, (tris, nrms) in {"cube.stl": cube_mesh(), "sphere.stl": sphere_mesh(2), "fan.stl": fan_mesh(),
                               "patch.stl": (patch, pn)}.items():
        write_stl(tmp_path / "res" / "stl" / name, file_axes(tris), file_axes(nrms))
        rp, rn, ri = R.stl_load(str(tmp_path / "work"), name)
        s = H.load_mesh(str(tmp_path / "res" / "stl" / name))
        assert s.pos.shape == rp.shape, name
        np.testing.assert_array_equal(s.pos[s.idx], rp[ri], err_msg=name)
        np.testing.assert_array_equal(s.nrm[s.idx], rn[ri], err_msg=name)


def test_scene_file_with_stl_mesh(tmp_path):
    """A scene naming an .stl mesh reads it from the stl directory beside the
    mesh directory (the reference's res/obj and res/stl)."""
    (tmp_path / "obj").mkdir()
    (tmp_path / "stl").mkdir()
    (tmp_path / "scene").mkdir()
    tris, nrms = cube_mesh()
    write_stl(tmp_path / "stl" / "cube.stl", file_axes(tris), file_axes(nrms))
    (tmp_path / "scene" / "s.scene").write_text("cube.stl\n\tposition(1, 2, 3)\n\tscale(2, 2, 2)\n")
    s = H.load_scene(str(tmp_path / "scene" / "s.scene"), str(tmp_path / "obj"))
    pos, _ = corner_data(s)
    np.testing.assert_allclose(pos, tris * 2 + np.float32([1, 2, 3]), atol=1e-6)
    (tmp_path / "scene" / "m.scene").write_text("gone.stl\n\tposition(0, 0, 0)\n")
    with pytest.raises(RTError, match="Could not load STL"):
        H.load_scene(str(tmp_path / "scene" / "m.scene"), str(tmp_path / "obj"))


def test_ppm_and_png_round_trips(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (17, 23, 3), dtype=np.uint8)
    I.write_ppm(tmp_path / "a.ppm", img)
    assert np.array_equal(I.read_ppm(tmp_path / "a.ppm"), img)
    assert (tmp_path / "a.ppm").read_bytes().startswith(b"P6\n23\n17\n255\n")
    I.ppm_to_png(tmp_path / "a.ppm", tmp_path / "a.png")
    with Image.open(tmp_path / "a.png") as im:
        assert im.mode == "RGB" and np.array_equal(np.array(im), img)
    # a PPM another writer made (comment, single-space header) reads the same
    Image.fromarray(img).save(tmp_path / "b.ppm")
    assert np.array_equal(I.read_ppm(tmp_path / "b.ppm"), img)


def test_to_rgb8_truncates():
    v = np.array([[-1.0, 0.0, 0.5], [0.999, 1.0, 7.0]], np.float32)
    assert I.to_rgb8(v).tolist() == [[0, 0, 127], [254, 255, 255]]


def test_compare_metric_closed_forms(tmp_path):
    """Identical images: 0.  One channel of one pixel off by d: (d / 255)^2
    * 255 / 3 over width * height.  The difference image as PIL stores the
    reference's putdata (values clipped to 255)."""
    from PIL import Image
    a = np.full((4, 5, 3), 30, np.uint8)
    m, diff = I.compare(a, a)
    assert m == 0.0 and not diff.any()
    b = a.copy()
    b[2, 3, 1] = 30 + 200
    m, diff = I.compare(a, b, diff_path=tmp_path / "d.png")
    e = (200 / 255.0) ** 2 * 255.0 / 3.0
    assert abs(m - e / 20) < 1e-12
    v = int(e) * 3
    assert diff[2, 3].tolist() == [min(v, 255), v // 2, min(v, 255)]
    assert m == I.compare(a, b, size=(5, 4))[0]
    assert abs(I.compare(a, b, size=(1920, 1080))[0] - e / (1920 * 1080)) < 1e-15
    # difference image: the same values PIL's putdata keeps for the reference's tuples
    rng = np.random.default_rng(9)
    x = rng.integers(0, 256, (6, 7, 3), dtype=np.uint8)
    y = rng.integers(0, 256, (6, 7, 3), dtype=np.uint8)
    _, diff = I.compare(x, y)
    data = []
    for r1, r2 in zip(x, y):
        for i1, i2 in zip(r1, r2):
            err = sum((abs(int(i1[c]) - int(i2[c])) / 255.0) ** 2 * 255.0 for c in range(3)) / 3.0
            err = int(err) * 3
            data.append((err, int(err / 2), err))
    im = Image.new("RGB", (7, 6))
    im.putdata(data)
    assert np.array_equal(np.array(im), diff)
    with Image.open(tmp_path / "d.png") as im2:
        assert np.array(im2).shape == (4, 5, 3)
