"""Builds libffcm_hip.so (the gfx950 backend) in place, next to this file.

Every translation unit under csrc/ is compiled by hipcc for gfx950 only and linked into one
shared library with a plain C ABI (include/ffcm_hip.h). Packed FP32 instructions are disabled
for all device code.

    python ffcm.jl_amd/build.py            # build
    python ffcm.jl_amd/build.py --clean    # remove the objects and the library
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(ROOT, "_build", "ffcm_hip")
LIB = os.path.join(HERE, "libffcm_hip.so")
HOST_LIB = os.path.join(HERE, "libffcm_suspension.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["common.hip", "cell_list.hip", "spread.hip", "fft.hip", "interp.hip", "correct.hip", "adjoint.hip", "api.cpp", "batch.cpp",
           "noalloc.hip", "seam.cpp",
           "fft_plans_0.hip", "fft_plans_1.hip", "fft_plans_2.hip", "fft_plans_3.hip"]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    host_src = os.path.join(HERE, "csrc_host", "suspension.cpp")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-o", HOST_LIB + ".tmp", host_src],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on suspension.cpp:\n{r.stderr}")
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)
    for f in (LIB, HOST_LIB):
        if os.path.exists(f):
            os.remove(f)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    else:
        build(verbose=True)
