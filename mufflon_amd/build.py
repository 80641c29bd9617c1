"""Build of the HIP library (gfx950) in-tree with hipcc.

Every source is compiled with -ffp-contract=off (no implicit fused
multiply-add: float results are fixed by the source), with packed FP32
instructions disabled, and linked into mufflon_amd/lib/libmufflon_hip.so.
The scene-file loader (host code only, zlib) is linked against it into
mufflon_amd/lib/libmffloader.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_obj")
LIB = os.path.join(PKG, "lib")
LIBNAME = os.path.join(LIB, "libmufflon_hip.so")
LOADERNAME = os.path.join(LIB, "libmffloader.so")
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MFH_OFFLOAD_ARCH", "gfx950")

SOURCES = ["sort_scan.hip", "lbvh.hip", "trace.hip", "pt.hip", "pt_wavefront.hip", "pt_adjoint.hip", "wireframe.hip", "lt.hip",
           "capi.cpp",
           "light_tree.cpp", "pt_host.cpp", "core_api.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
DEVICE = [
    "-x", "hip", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + COMMON + (DEVICE if src.endswith(".hip") else []) + ["-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return out


def build_library(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIBNAME + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIBNAME)
    if verbose:
        print("built", LIBNAME)
    build_loader(verbose)
    return LIBNAME


def build_loader(verbose=False):
    tmp = LOADERNAME + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(CSRC, "mff_loader.cpp"), "-o", tmp, "-L" + LIB, "-lmufflon_hip", "-lz",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("loader build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LOADERNAME)
    if verbose:
        print("built", LOADERNAME)
    return LOADERNAME


if __name__ == "__main__":
    build_library(verbose=True)
    sys.exit(0)
