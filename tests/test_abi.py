"""The C-ABI libraries load and export every function their headers declare;
argument validation and error reporting work without a GPU (no compute
calls are made here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from master_thesis_amd import _lib as L
from master_thesis_amd import host_lib as H

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: str) -> set[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b((?:rt|mth)_[a-z0-9_]+)\s*\(", text))


def _exported(lib: Path) -> set[str]:
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_librt_exports_every_declared_symbol():
    decl = _declared("rt.h")
    assert decl, "no declarations parsed"
    missing = decl - _exported(L.LIB_PATH)
    assert not missing, missing
    assert decl == set(L.SIGNATURES), decl ^ set(L.SIGNATURES)


def test_librt_host_exports_every_declared_symbol():
    decl = _declared("rt_host.h")
    missing = decl - _exported(H.LIB_PATH)
    assert not missing, missing


def test_librt_loads_and_binds():
    lib = L.load()
    for name in L.SIGNATURES:
        assert hasattr(lib, name)


def test_null_context_is_rejected():
    lib = L.load()
    assert lib.rt_scene_upload(None, None) == L.RT_E_INVALID
    assert lib.rt_build(None, None, None) == L.RT_E_INVALID
    assert lib.rt_trace_closest(None, None, 0, None, 0, None) == L.RT_E_INVALID
    assert lib.rt_render(None, None, None, None, None, 0) == L.RT_E_INVALID
    assert lib.rt_get_timing(None, None) == L.RT_E_INVALID
    lib.rt_destroy(None)  # no-op


def test_create_without_out_pointer():
    lib = L.load()
    assert lib.rt_create(None, 0) == L.RT_E_INVALID
    assert b"NULL" in lib.rt_last_error(None)


def test_header_constants_match_binding():
    """Every #define of include/rt.h that the Python binding mirrors has the
    same value there."""
    import re
    text = (ROOT / "include" / "rt.h").read_text()
    found = 0
    for name, expr in re.findall(r"^#define\s+(RT_\w+)\s+(.+?)\s*(?:/\*.*)?$", text, re.M):
        if not hasattr(L, name):
            continue
        value = eval(re.sub(r"\b(0x[0-9A-Fa-f]+|\d+)u\b", r"\1", expr), {})
        assert getattr(L, name) == value, (name, expr, getattr(L, name))
        found += 1
    assert found >= 10


def test_struct_layouts_match_header():
    # sizes are asserted at import; offsets of the fields the kernels read
    assert L.rt_ray.dx.offset == 16 and L.rt_ray.tmax.offset == 28
    assert L.rt_hit.t.offset == 4
    assert L.rt_material.light_source.offset == 28
    assert L.rt_bvh4_node.child.offset == 96
    assert L.rt_tri_rec.e1.offset == 32
    assert L.rt_camera.origin.offset == 128 and L.rt_camera.width.offset == 140
    assert L.rt_settings.seed.offset == 8


def test_host_errors_are_reported(tmp_path):
    with pytest.raises(L.RTError, match="failed to open"):
        H.load_scene(str(tmp_path / "missing.scene"), str(tmp_path))
    with pytest.raises(L.RTError, match="fov"):
        H.camera((0, 0, 0), 1.0, 0.0, 0.0, 8, 8)


def test_missing_library_fails_loudly(tmp_path):
    """Without the built HIP library the product path raises at once; it has
    no CPU fallback to take instead."""
    import os
    import subprocess
    import sys
    env = dict(os.environ, MTRT_LIB=str(tmp_path / "absent" / "libmtrt.so"))
    code = "from master_thesis_amd.tracer import Tracer\nTracer(0)\n"
    p = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode != 0
    assert "is not built" in p.stderr
