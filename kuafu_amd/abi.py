"""ctypes mirror of include/kuafu_amd.h and the loader of libkuafu_amd.so.

The structures reproduce the reference's std430 / host layouts byte for byte
(vertex.hpp:28-34, geometry.hpp:70-99, light.hpp:41-58, camera.hpp:181-192,
rt.hpp:31-42).  Loading fails loudly when the HIP library is missing: there is
no CPU fallback for the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkuafu_amd.so")

# ----------------------------------------------------------------- layouts
VERTEX_DTYPE = np.dtype([("pos", "<f4", 3), ("normal", "<f4", 3), ("color", "<f4", 3),
                         ("uv", "<f4", 2), ("pad", "<f4")])
MATERIAL_DTYPE = np.dtype([("diffuse", "<f4", 4), ("emission", "<f4", 4), ("alpha", "<f4"),
                           ("metallic", "<f4"), ("specular", "<f4"), ("roughness", "<f4"),
                           ("ior", "<f4"), ("transmission", "<f4"), ("diffuse_tex", "<i4"),
                           ("metallic_tex", "<i4"), ("roughness_tex", "<i4"),
                           ("transmission_tex", "<i4"), ("pad0", "<i4"), ("pad1", "<i4")])
INSTANCE_DTYPE = np.dtype([("transform", "<f4", 16), ("geometry_index", "<u4"), ("pad", "<u4", 3)])
RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("inst", "<u4"), ("prim", "<u4"), ("b1", "<f4"), ("b2", "<f4"),
                      ("back", "<u4")])
assert VERTEX_DTYPE.itemsize == 48 and MATERIAL_DTYPE.itemsize == 80 and INSTANCE_DTYPE.itemsize == 80
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 24

KF_GEOM_OPAQUE = 1
KF_GEOM_HIDDEN = 2
KF_ENV_CLEAR_COLOR = 0
KF_ENV_CUBEMAP = 1
KF_ENV_GRADIENT = 2
MAX_POINT_LIGHTS = 32
MAX_ACTIVE_LIGHTS = 8


class DirLight(C.Structure):
    _fields_ = [("direction", C.c_float * 4), ("rgbs", C.c_float * 4)]


class PointLights(C.Structure):
    _fields_ = [("posr", (C.c_float * 4) * MAX_POINT_LIGHTS), ("rgbs", (C.c_float * 4) * MAX_POINT_LIGHTS)]


class ActiveLights(C.Structure):
    _fields_ = [("view", (C.c_float * 16) * MAX_ACTIVE_LIGHTS), ("proj", (C.c_float * 16) * MAX_ACTIVE_LIGHTS),
                ("front", (C.c_float * 4) * MAX_ACTIVE_LIGHTS), ("rgbs", (C.c_float * 4) * MAX_ACTIVE_LIGHTS),
                ("position", (C.c_float * 4) * MAX_ACTIVE_LIGHTS), ("sftp", (C.c_float * 4) * MAX_ACTIVE_LIGHTS)]


class CameraUBO(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("view_inverse", C.c_float * 16),
                ("proj_inverse", C.c_float * 16), ("position", C.c_float * 4), ("front", C.c_float * 4),
                ("pad1", C.c_float * 4), ("pad2", C.c_float * 4)]


class PushConstants(C.Structure):
    _fields_ = [("clear_color", C.c_float * 4), ("frame_count", C.c_int32), ("spp", C.c_uint32),
                ("path_depth", C.c_uint32), ("use_env_map", C.c_uint32), ("russian_roulette", C.c_uint32),
                ("rr_min_bounces", C.c_uint32), ("nee", C.c_uint32), ("nee_min_bounces", C.c_uint32)]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba8", C.c_void_p)]


class SceneDesc(C.Structure):
    _fields_ = [("num_geometries", C.c_uint32), ("geom_vertex_offset", C.c_void_p),
                ("geom_tri_offset", C.c_void_p), ("geom_flags", C.c_void_p), ("vertices", C.c_void_p),
                ("indices", C.c_void_p), ("mat_index", C.c_void_p), ("num_instances", C.c_uint32),
                ("instances", C.c_void_p), ("num_materials", C.c_uint32), ("materials", C.c_void_p),
                ("num_textures", C.c_uint32), ("textures", C.c_void_p), ("dir_light", C.c_void_p),
                ("point_lights", C.c_void_p), ("active_lights", C.c_void_p), ("env_kind", C.c_uint32),
                ("env_faces", C.c_void_p), ("env_gradient", C.c_float * 8), ("num_envs", C.c_uint32),
                ("env_instance_offset", C.c_void_p), ("separate_geometry", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64), ("trace_ms", C.c_float),
                ("extend_ms", C.c_float), ("shadow_ms", C.c_float), ("other_ms", C.c_float),
                ("extend_launches", C.c_uint32), ("shadow_launches", C.c_uint32),
                ("instance_visits", C.c_uint64), ("speculative_rays", C.c_uint64),
                ("tlas_node_visits", C.c_uint64), ("tail_rays", C.c_uint64), ("graph_captures", C.c_uint64)]

    def rays(self):
        return int(self.primary_rays + self.extension_rays + self.shadow_rays)


class TuneResult(C.Structure):
    _fields_ = [("samples_per_batch", C.c_uint32), ("refill", C.c_uint32), ("grab", C.c_uint32),
                ("frame_ms", C.c_float), ("default_ms", C.c_float), ("trials", C.c_uint32),
                ("tail_bounce", C.c_uint32), ("grab0", C.c_uint32), ("grab_shadow", C.c_uint32),
                ("tail_pairs", C.c_uint32)]


class RenderOpts(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("samples_per_batch", C.c_uint32), ("keep_samples", C.c_uint32),
                ("count_traversal", C.c_uint32), ("use_graph", C.c_uint32), ("full_tlas_camera", C.c_uint32)]


class BvhInfo(C.Structure):
    _fields_ = [("blas_nodes", C.c_uint32), ("tlas_nodes", C.c_uint32), ("triangles", C.c_uint32),
                ("instances", C.c_uint32), ("max_blas_depth", C.c_uint32), ("max_tlas_depth", C.c_uint32),
                ("build_ms", C.c_float), ("blas4_nodes", C.c_uint32), ("tlas4_nodes", C.c_uint32)]


assert C.sizeof(DirLight) == 32 and C.sizeof(PointLights) == 1024 and C.sizeof(ActiveLights) == 1536
assert C.sizeof(CameraUBO) == 320 and C.sizeof(PushConstants) == 48


def ptr(a):
    """Address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


# ----------------------------------------------------------------- library
# (name, restype, argtypes) of every entry point declared in include/kuafu_amd.h
EXPORTS = [
    ("kf_init", C.c_int, [C.c_int]),
    ("kf_last_error", C.c_char_p, []),
    ("kf_version", C.c_char_p, []),
    ("kf_camera_ubo_compute", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float,
                                        C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p]),
    ("kf_world_to_object", C.c_int, [C.c_void_p, C.c_void_p]),
    ("kf_active_light_fill", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float,
                                       C.c_int]),
    ("kf_scene_create", C.c_int, [C.c_void_p, C.c_void_p]),
    ("kf_scene_destroy", None, [C.c_void_p]),
    ("kf_scene_update_instances", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("kf_scene_update_geometry", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
    ("kf_ktx_cubemap_read", C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("kf_scene_update_materials", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("kf_scene_set_lights", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("kf_scene_bvh_info", C.c_int, [C.c_void_p, C.c_void_p]),
    ("kf_scene_read_blas", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_uint32]),
    ("kf_scene_read_bvh4", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_uint32]),
    ("kf_trace_closest", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("kf_trace_any", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("kf_trace_closest_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("kf_trace_closest_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("kf_renderer_create", C.c_int, [C.c_uint32, C.c_uint32, C.c_void_p]),
    ("kf_renderer_destroy", None, [C.c_void_p]),
    ("kf_render", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("kf_render_envs", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p]),
    ("kf_renderer_autotune", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("kf_renderer_set_schedule", C.c_int, [C.c_void_p] + [C.c_uint32] * 6),
    ("kf_renderer_set_tail_pairs", C.c_int, [C.c_void_p, C.c_uint32]),
    ("kf_renderer_read", C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ("kf_renderer_write", C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ("kf_renderer_read_samples", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("kf_renderer_device_target", C.c_void_p, [C.c_void_p, C.c_int]),
    ("kf_renderer_download_bgra8", C.c_int, [C.c_void_p, C.c_void_p]),
    ("kf_renderer_download_gray32f", C.c_int, [C.c_void_p, C.c_void_p]),
    ("kf_bench_copy", C.c_int, [C.c_uint64, C.c_int, C.c_void_p]),
    ("kf_denoiser_create", C.c_int, [C.c_char_p, C.c_void_p]),
    ("kf_denoiser_destroy", None, [C.c_void_p]),
    ("kf_denoiser_init_random", C.c_int, [C.c_void_p, C.c_uint32]),
    ("kf_denoiser_save", C.c_int, [C.c_void_p, C.c_char_p]),
    ("kf_denoiser_load", C.c_int, [C.c_void_p, C.c_char_p]),
    ("kf_denoiser_num_params", C.c_int, [C.c_void_p, C.c_void_p]),
    ("kf_denoiser_get_params", C.c_int, [C.c_void_p, C.c_void_p]),
    ("kf_denoiser_set_params", C.c_int, [C.c_void_p, C.c_void_p]),
    ("kf_denoiser_apply", C.c_int, [C.c_void_p, C.c_void_p]),
    ("kf_denoiser_infer_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_void_p]),
    ("kf_denoiser_infer_device_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                 C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("kf_denoiser_train_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_float, C.c_void_p]),
    ("kf_denoiser_gradients", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_void_p, C.c_void_p]),
    ("kf_denoiser_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    ("kf_conv3x3_forward", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                     C.c_uint32, C.c_uint32, C.c_int, C.c_void_p]),
    ("kf_conv3x3_backward", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                      C.c_uint32, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
]

_lib = None


def lib():
    """Load libkuafu_amd.so (built by kuafu_amd/build.py); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libkuafu_amd.so is not built (run __graft_entry__.build()): %s" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, res, args in EXPORTS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class KuafuError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise KuafuError(lib().kf_last_error().decode())
    return rc
