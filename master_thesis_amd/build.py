"""Builds the native libraries in-tree.

  master_thesis_amd/libmtrt.so    HIP kernels + the C ABI of include/rt.h (hipcc, gfx950)
  master_thesis_amd/libmtrt_host.so C++ scene / camera / config layer (g++), see host/
  master_thesis_amd/sw_render     SoftwareRenderer-compatible command line (g++ + libmtrt.so)

Every HIP source is compiled with -ffp-contract=off and IEEE division/sqrt:
the hit and path arithmetic must round exactly like the CPU restatement in
oracle/ for the results to match bit for bit.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # no packed-FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-I", str(ROOT / "include")]

LIBRT = PKG / "libmtrt.so"
LIBHOST = PKG / "libmtrt_host.so"
CLI = PKG / "sw_render"


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the device-only feature flag; that warning is harmless
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build_librt(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [BUILD / (s.stem + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)]) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = LIBRT.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIBRT)
    return LIBRT


def build_host() -> list[Path]:
    # libmtrt_host.so has no GPU dependency (formats, transforms, camera);
    # the engine (which calls libmtrt) is linked only into the command line
    gpu_srcs = {"engine.cpp", "sw_render_main.cpp"}
    lib_srcs = [s for s in sorted(HOST.glob("*.cpp")) if s.name not in gpu_srcs]
    out = []
    tmp = LIBHOST.with_suffix(".so.tmp")
    _run([CXX, *CXX_FLAGS, "-shared", "-o", str(tmp), *map(str, lib_srcs), "-pthread"])
    os.replace(tmp, LIBHOST)
    out.append(LIBHOST)
    tmp = CLI.with_suffix(".tmp")
    _run([CXX, *CXX_FLAGS, "-o", str(tmp), str(HOST / "sw_render_main.cpp"), str(HOST / "engine.cpp"),
          *map(str, lib_srcs), "-L", str(PKG), "-lmtrt", "-Wl,-rpath,$ORIGIN", "-pthread"])
    os.replace(tmp, CLI)
    out.append(CLI)
    return out


def build_all() -> None:
    build_librt()
    build_host()


if __name__ == "__main__":
    build_all()
    print("built", LIBRT)
