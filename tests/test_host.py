"""Host layer (libmtrt_host.so) and the sw_render command line: the
reference's text formats, object transforms, camera and PPM output.  The
scene, mesh, camera and renderer files are written here, fresh."""
import math
import subprocess

import numpy as np
import pytest

from master_thesis_amd import build as B
from master_thesis_amd import host_lib as H
from master_thesis_amd._lib import RTError

PANEL_OBJ = """# a 2x2 quad in the x = 0 plane, normal +x
v 0.0 1.0 -1.0
v 0.0 -1.0 -1.0
v 0.0 1.0 1.0
v 0.0 -1.0 1.0
vt 0.0 0.0
vn 1.0 0.0 0.0
f 2/1/1 1/1/1 3/1/1 4/1/1
"""

TRI_OBJ = """v 0 0 0
v 1 0 0
v 0 1 0
vn 0 0 1
f 1//1 2//1 3//1
"""


def _files(tmp_path, scene_text, objs):
    (tmp_path / "obj").mkdir()
    (tmp_path / "scene").mkdir()
    for name, text in objs.items():
        (tmp_path / "obj" / name).write_text(text)
    sp = tmp_path / "scene" / "s.scene"
    sp.write_text(scene_text)
    return sp, tmp_path / "obj"


def test_obj_fan_triangulation_and_unique_vertices(tmp_path):
    sp, od = _files(tmp_path, "panel.obj\n\tposition(0, 0, 0)\n", {"panel.obj": PANEL_OBJ})
    s = H.load_scene(str(sp), str(od))
    assert s.idx.shape == (2, 3) and s.pos.shape == (4, 3)
    # face 2 1 3 4 -> (2,1,3), (2,3,4) in first-seen vertex numbering 0..3
    assert s.idx.tolist() == [[0, 1, 2], [0, 2, 3]]
    assert np.array_equal(s.pos[0], [0.0, -1.0, -1.0])
    assert np.all(s.nrm == np.array([1, 0, 0], np.float32))


def test_scene_keys_and_thresholds(tmp_path):
    text = ("# comment line\n"
            "tri.obj\n\tposition(1, 2, 3)\n\tcolor(0.5, 0.25, 1)\n\tdiffuse(2)\n\treflect(1)\n\ttransparent(1)\n"
            "\trefractionIndex(1.5)\n\n"
            "Probes\n\tprobeCount(2, 2, 2)\n\tprobeStartCorner(-1, -1, -1)\n\tbetweenProbeDistance(2, 2, 2)\n\n"
            "tri.obj\n\tposition(0, 0, 0)\n\tlightSource(1.0)\n\tcolor(15, 15, 15)\n")
    sp, od = _files(tmp_path, text, {"tri.obj": TRI_OBJ})
    s = H.load_scene(str(sp), str(od))
    assert s.materials.shape == (2,)
    m = s.materials[0]
    assert tuple(m["color"]) == (0.5, 0.25, 1.0)
    assert m["diffuse_threshold"] == np.float32(0.5) and m["reflect_threshold"] == np.float32(0.75)
    assert m["transparent_threshold"] == np.float32(1.0) and m["refraction_index"] == np.float32(1.5)
    assert s.materials[1]["light_source"] == 1 and s.materials[0]["light_source"] == 0
    assert np.array_equal(s.pos[:3], np.array([[1, 2, 3], [2, 2, 3], [1, 3, 3]], np.float32))
    assert s.tri_mat.tolist() == [0, 1]


def test_object_transform_scale_rotate_translate(tmp_path):
    """v' = v * S * R(axis, angle) * T in float; checked against float64 maths."""
    text = "tri.obj\n\tposition(0.5, -1, 2)\n\tscale(2, 3, 4)\n\trotation(0, 1, 0, 0.7)\n"
    sp, od = _files(tmp_path, text, {"tri.obj": TRI_OBJ})
    s = H.load_scene(str(sp), str(od))
    a = 0.7
    # row-vector rotation about +y by the quaternion's matrix (rotation.cpp:35-56)
    R = np.array([[math.cos(a), 0, -math.sin(a)], [0, 1, 0], [math.sin(a), 0, math.cos(a)]])
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float) * [2, 3, 4]
    want = v @ R + [0.5, -1, 2]
    assert np.allclose(s.pos, want, atol=1e-6)
    # normals get the same matrix with w = 0 and are not renormalised
    assert np.allclose(s.nrm[0], np.array([0, 0, 4.0]) @ R, atol=1e-6)


def test_camera_looks_where_the_angles_say():
    # angle(1, 0): theta = pi, phi = pi/2 -> looking along -x
    cam = H.camera((4.0, 0.0, 0.0), 1.0, 0.0, math.pi / 4, 64, 64)
    import oracle as O
    d = O.primary_dir(cam, 32, 32)
    assert d[0] < -0.999 and abs(d[1]) < 0.02 and abs(d[2]) < 0.02
    assert np.allclose(np.array(cam.origin[:]), [4, 0, 0], atol=1e-6)
    # the first pixel column's ray makes half the field of view (in x) with the axis
    cam2 = H.camera((0.0, 0.0, 0.0), 1.0, 0.0, math.radians(40.0), 400, 200)
    left = O.primary_dir(cam2, 0, 100)
    half = math.atan2(abs(left[2]), -left[0])
    # pixel centre 0.5 px in from the edge: tan(angle) = aspect * tan(fov / 2) * (1 - 1/400)
    assert abs(math.tan(half) - 2.0 * math.tan(math.radians(20.0)) * (1 - 1 / 400)) < 1e-4


def test_renderer_file(tmp_path):
    p = tmp_path / "r.renderer"
    p.write_text("PathTracer\n\tvisionJumpCount(5)\n\traysPerPixel(2500)\n")
    name, s = H.renderer_from_file(str(p))
    assert name == "PathTracer" and s.max_depth == 5 and s.spp == 2500 and s.seed == 1


def test_camera_file_with_fov_extension(tmp_path):
    p = tmp_path / "c.camera"
    p.write_text("Camera\n\tposition(4, 0, 0)\n#\ttheta, phi\n\tangle(1, 0)\n\tfov(40)\n")
    a = H.camera_from_file(str(p), 32, 32)
    b = H.camera((4.0, 0.0, 0.0), 1.0, 0.0, float(np.float32(math.radians(40.0))), 32, 32)
    assert np.array_equal(np.array(a.proj_inv[:]), np.array(b.proj_inv[:]))


def test_ppm_writer(tmp_path):
    img = np.arange(2 * 3 * 3, dtype=np.uint8).reshape(2, 3, 3)
    p = tmp_path / "o.ppm"
    H.write_ppm(str(p), img)
    data = p.read_bytes()
    assert data.startswith(b"P6\n3\n2\n255\n") and data[len(b"P6\n3\n2\n255\n"):] == img.tobytes()


def test_missing_mesh_and_bad_index(tmp_path):
    sp, od = _files(tmp_path, "nothere.obj\n\tposition(0, 0, 0)\n", {})
    with pytest.raises(RTError, match="Could not load OBJ"):
        H.load_scene(str(sp), str(od))
    bad = "v 0 0 0\nvn 0 0 1\nf 1//1 2//1 3//1\n"
    sp2 = tmp_path / "b.scene"
    sp2.write_text("bad.obj\n\tposition(0, 0, 0)\n")
    (od / "bad.obj").write_text(bad)
    with pytest.raises(RTError, match="out of range"):
        H.load_scene(str(sp2), str(od))


def test_cli_argument_count():
    p = subprocess.run([str(B.CLI)], capture_output=True, text=True)
    assert p.returncode == 255 and "wrong paramter count" in p.stdout


def test_cli_unknown_renderer(tmp_path):
    sp, od = _files(tmp_path, "tri.obj\n\tposition(0, 0, 0)\n", {"tri.obj": TRI_OBJ})
    r = tmp_path / "x.renderer"
    r.write_text("PhotonMapper\n\tvisionJumpCount(5)\n\traysPerPixel(1)\n")  # a Vulkan-only renderer
    c = tmp_path / "c.camera"
    c.write_text("Camera\n\tposition(4, 0, 0)\n\tangle(1, 0)\n")
    p = subprocess.run([str(B.CLI), str(r), str(sp), "8", "8", "1", str(c), str(tmp_path / "o.ppm")],
                       capture_output=True, text=True)
    assert p.returncode == 1 and "getRenderer not found" in p.stderr
