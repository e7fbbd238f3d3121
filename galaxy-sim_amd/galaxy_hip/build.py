"""Compiles libgalaxy_hip.so (HIP, gfx950) and the N-API addon in-tree.

  python galaxy-sim_amd/galaxy_hip/build.py          # library + addon
Outputs:
  galaxy-sim_amd/lib/libgalaxy_hip.so
  galaxy-sim_amd/node/galaxy_hip.node   (when node_api.h is present)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "lib", "obj")
LIB = os.path.join(LIBDIR, "libgalaxy_hip.so")
ADDON_SRC = os.path.join(PKG, "node", "addon.cc")
ADDON = os.path.join(PKG, "node", "galaxy_hip.node")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["encode.hip", "sort.hip", "build.hip", "refresh.hip", "walk.hip", "records.hip",
           "stellar.hip", "api.hip", "noalloc.hip"]

# -ffp-contract=off: every a*b+c in the kernels rounds twice, as the reference
# expressions do (the walk's opening test and the FP64 COM sums depend on it).
# Packed-FP32 instructions are disabled for the whole library.
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-gpu-rdc", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r


def build_library(jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)

    def one(src):
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_addon():
    """N-API addon (node_api.h, N-API v8) linked against libgalaxy_hip.so."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")) or not os.path.exists(ADDON_SRC):
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = ADDON + ".tmp"
    _run([cxx, "-O2", "-fPIC", "-shared", "-std=c++17", "-DNAPI_VERSION=8", "-DNODE_GYP_MODULE_NAME=galaxy_hip", "-Wall",
          "-I", inc,
          "-I", os.path.join(ROOT, "include"), ADDON_SRC, "-o", tmp,
          "-L", LIBDIR, "-lgalaxy_hip", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, ADDON)
    return ADDON


def build_all():
    lib = build_library()
    addon = build_addon()
    return lib, addon


if __name__ == "__main__":
    print(build_all())
