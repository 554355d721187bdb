"""Builds libradgpu.so (HIP for gfx950 + host C++) in-tree.

    python -m pyradiance_amd.build

Objects are compiled in parallel into pyradiance_amd/_build/ and linked into
pyradiance_amd/libradgpu.so.  Every call recompiles.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libradgpu.so"
BUILD = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RADGPU_ARCH", "gfx950")

SOURCES = ["scene_io.cpp", "rg_scene.cpp", "rg_bvh.cpp", "rg_capi.cpp", "rg_render.hip", "rg_build.hip",
           "rg_dcsky.hip", "rg_sky.hip", "rg_view.hip", "rg_format.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-but-set-variable"]
# Device code: gfx950 only; packed-FP32 instruction selection disabled.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src: str) -> Path:
    obj = BUILD / (src + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".hip") or src == "rg_capi.cpp":
        cmd += DEVICE
        cmd += ["-x", "hip"]
    else:
        cmd += ["-x", "c++", "-D__HIP_PLATFORM_AMD__"]
    cmd += ["-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libradgpu.so failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
