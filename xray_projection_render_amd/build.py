"""Builds libxrp.so (HIP kernels for gfx950 + C++ host code) in-tree.

    python -m xray_projection_render_amd.build

Every translation unit is compiled with -ffp-contract=off: the FP64 camera,
ray and sample arithmetic must round exactly as the reference's unfused
Go code does, and FP32 multiply-adds in the kernels are explicit fmaf().
Packed FP32 instructions are disabled for device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxrp.so")
ARCH = os.environ.get("XRP_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-I" + os.path.join(HERE, "..", "include")]
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

SOURCES = [
    ("march.hip", True),
    ("voxel.hip", True),
    ("batch.hip", True),
    ("surrogate.hip", True),
    ("noalloc.hip", True),
    ("api.cpp", False),
    ("scene.cpp", False),
    ("camera.cpp", False),
    ("host_driver.cpp", False),
]


def _compile(src: str, device: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-c"] + (DEVICE if device else []) + \
        [os.path.join(CSRC, src), "-o", obj] + COMMON
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [(s, d) for s, d in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda sd: _compile(*sd), srcs))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "-shared", "-o", tmp] + objs +
                   ["--offload-arch=" + ARCH, "-lpthread", "-lz"], check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
