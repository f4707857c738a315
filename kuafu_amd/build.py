"""Build libkuafu_amd.so (HIP, gfx950) in-tree.

Every HIP translation unit is compiled with hipcc for gfx950 with
-ffp-contract=off (the kernels' operation order is part of their
specification) and packed FP32 disabled, then linked into one shared
library next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkuafu_amd.so")
FACADE_LIB = os.path.join(HERE, "libkuafu.so")
FACADE = ["glm_impl.cpp", "config_global.cpp", "geometry.cpp", "camera_scene.cpp", "kuafu.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["capi.cpp", "ktx.cpp", "bvh.hip", "trace.hip", "pathtrace.hip", "denoise.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-lambda-capture",
]


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", path, "-o", out] + FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, r.stdout, r.stderr))
    return out


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=gfx950", "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_facade()
    if verbose:
        print("built", LIB, "and", FACADE_LIB)
    return LIB


def build_facade():
    """libkuafu.so: the kuafu:: C++ API (include/kuafu.hpp) over the C ABI."""
    objs = []
    inc = os.path.join(os.path.dirname(HERE), "include")

    def one(src):
        out = os.path.join(OBJ, "facade_" + src + ".o")
        cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-I", inc, "-c",
               os.path.join(HERE, "facade", src), "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("g++ failed for %s:\n%s\n%s" % (src, r.stdout, r.stderr))
        return out

    with ThreadPoolExecutor(max_workers=len(FACADE)) as ex:
        objs = list(ex.map(one, FACADE))
    tmp = FACADE_LIB + ".tmp"
    cmd = ["g++", "-shared", "-o", tmp] + objs + ["-L", HERE, "-lkuafu_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, FACADE_LIB)
    build_facade_test()


def build_facade_test():
    """tests/cpp/facade_test: C++ client of the facade used by tests/test_facade.py."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "facade_test.cpp")
    if not os.path.exists(src):
        return
    out = os.path.join(root, "tests", "cpp", "facade_test")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(root, "include"), src, "-o", out,
           "-L", HERE, "-lkuafu", "-lkuafu_amd", "-Wl,-rpath," + HERE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade test build failed:\n%s\n%s" % (r.stdout, r.stderr))


if __name__ == "__main__":
    build()
    sys.exit(0)
