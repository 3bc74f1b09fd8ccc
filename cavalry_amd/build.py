"""Compile the HIP sources into cavalry_amd/_lib/libcavalry_amd.so for gfx950.

Run as ``python -m cavalry_amd.build``. Every source is compiled with
-ffp-contract=off and IEEE division/sqrt so the kernels reproduce the
reference's float arithmetic operation by operation.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libcavalry_amd.so")

SOURCES = [
    "radix_sort.hip",
    "lbvh.hip",
    "levels.hip",
    "bottomup.hip",
    "treelet.hip",
    "traverse.hip",
    "render.hip",
    "rlpath.hip",
    "queue_sort.hip",
    "serve.hip",
    "render_capi.cpp",
    "film_io.cpp",
    "image_read.cpp",
    "pbrt.cpp",
    "server.cpp",
    "capi.cpp",
    "noalloc.hip",
    "camera.cpp",
    "ctx_release.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    workers = min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the command-line renderer, linked against the library next to it
    cli = os.path.join(OUT_DIR, "cavalry")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(HERE, "host", "main.cpp"),
           "-I" + os.path.join(os.path.dirname(HERE), "include"), "-L" + OUT_DIR, "-lcavalry_amd",
           "-Wl,-rpath,$ORIGIN", "-o", cli + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(cli + ".tmp", cli)
    if verbose:
        print(f"built {LIB} and {cli}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
