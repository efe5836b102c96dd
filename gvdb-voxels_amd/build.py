"""Builds libgvdb_amd.so (HIP kernels for gfx950 + host runtime + C ABI).

Usage: python gvdb-voxels_amd/build.py   (or through __graft_entry__.build())
Every HIP source is compiled with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libgvdb_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GVDB_OFFLOAD_ARCH", "gfx950")

SOURCES = ["util.hip", "topology.hip", "points.hip", "apron.hip", "render.hip", "compute.hip", "pipeline.cpp", "batch.cpp", "serve.cpp", "volume.cpp", "vbx.cpp", "plugin.cpp", "camera.cpp", "capi.cpp", "tune.cpp", "graph.cpp", "keysort.hip", "keybuild.cpp"]
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-I" + SRC, "-I" + os.path.join(HERE, "..", "include")]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(SRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    # C++ facade sample linked against the library (runs on a GPU host)
    demo = os.path.join(LIBDIR, "facade_demo")
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "examples", "facade_demo.cpp"), "-o", demo + ".tmp", "-L" + LIBDIR, "-lgvdb_amd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade_demo build failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(demo + ".tmp", demo)
    # sample user kernels as a code object for gvdb_compute_kernel / gvdb_render_kernel
    hsaco = os.path.join(LIBDIR, "user_kernels.hsaco")
    cmd = [HIPCC, "--genco", "-O3", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "examples", "user_kernels.hip"), "-o", hsaco + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("user kernel build failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(hsaco + ".tmp", hsaco)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
