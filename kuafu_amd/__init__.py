"""kuafu_amd — MI355X (gfx950) path-tracing hot path of Kuafu.

The product is libkuafu_amd.so (HIP kernels + C ABI, include/kuafu_amd.h) and
the C++ facade (include/kuafu.hpp).  This package is the Python host used by
tests and the benchmark: ctypes bindings (abi), the scene API mirror (api) and
workload scenes (scenes).
"""
from .abi import KuafuError, lib  # noqa: F401
from .api import (GpuScene, Renderer, SceneData, camera_ubo, look_at, mat_identity, mat_rotate,  # noqa: F401
                  mat_scale, mat_translate, material, push_constants)
