"""Python mirror of the kuafu scene API over the C ABI (test and bench host).

SceneData collects geometries, materials, instances, textures, lights and the
environment in the reference's layouts (kuafu::Scene, scene.hpp:47-211;
NiceMaterial geometry.hpp:11-30; lights light.hpp:9-39).  GpuScene uploads it
and builds every BLAS and the TLAS on the GPU; Renderer runs frames.
"""
import ctypes as C

import numpy as np

from . import abi
from .abi import check, lib, ptr

# ------------------------------------------------------------------ glm-style math
# Matrices are stored like glm::mat4 in memory: m[c, r] (column-major), so
# m.ravel() is the float[16] the C ABI expects.


def mat_identity():
    return np.eye(4, dtype=np.float32)


def mat_mul(a, b):
    """glm a * b for column-major stored matrices."""
    return (b.astype(np.float64) @ a.astype(np.float64)).astype(np.float32)


def mat_translate(v, m=None):
    t = mat_identity()
    t[3, :3] = v
    return t if m is None else mat_mul(m, t)


def mat_scale(s, m=None):
    s = np.broadcast_to(np.asarray(s, np.float32), (3,))
    t = mat_identity()
    t[0, 0], t[1, 1], t[2, 2] = s
    return t if m is None else mat_mul(m, t)


def mat_rotate(angle, axis, m=None):
    """glm::rotate(m, angle, axis)."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    c, s = np.cos(angle), np.sin(angle)
    x, y, z = a
    R = np.array([[c + (1 - c) * x * x, (1 - c) * x * y - s * z, (1 - c) * x * z + s * y],
                  [(1 - c) * y * x + s * z, c + (1 - c) * y * y, (1 - c) * y * z - s * x],
                  [(1 - c) * z * x - s * y, (1 - c) * z * y + s * x, c + (1 - c) * z * z]])
    t = mat_identity()
    t[:3, :3] = R.T  # stored column-major
    return t if m is None else mat_mul(m, t)


def mat_from_rows(R, t):
    """Affine transform from a 3x3 rotation/scale (math layout) and translation."""
    m = mat_identity()
    m[:3, :3] = np.asarray(R, np.float64).T
    m[3, :3] = t
    return m


def look_at(eye, center, up):
    """glm::lookAt (right-handed), stored column-major."""
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, 0], m[1, 0], m[2, 0] = s
    m[0, 1], m[1, 1], m[2, 1] = u
    m[0, 2], m[1, 2], m[2, 2] = -f
    m[3, 0] = -s @ eye
    m[3, 1] = -u @ eye
    m[3, 2] = f @ eye
    return m.astype(np.float32)


# ------------------------------------------------------------------ materials
def material(diffuse=(1.0, 1.0, 1.0), alpha=1.0, metallic=0.0, specular=0.5, roughness=0.5, ior=1.4,
             transmission=0.0, emission=(1.0, 1.0, 1.0), emission_strength=0.0, diffuse_tex=-1,
             metallic_tex=-1, roughness_tex=-1, transmission_tex=-1):
    """NiceMaterial (geometry.hpp:11-30) -> NiceMaterialSSBO (scene.cpp:330-339)."""
    m = np.zeros((), abi.MATERIAL_DTYPE)
    m["diffuse"] = (*diffuse, 0.0)
    m["emission"] = (*emission, emission_strength)
    m["alpha"], m["metallic"], m["specular"], m["roughness"] = alpha, metallic, specular, roughness
    m["ior"], m["transmission"] = ior, transmission
    m["diffuse_tex"], m["metallic_tex"] = diffuse_tex, metallic_tex
    m["roughness_tex"], m["transmission_tex"] = roughness_tex, transmission_tex
    return m


def vertices(pos, normal=None, uv=None):
    n = len(pos)
    v = np.zeros(n, abi.VERTEX_DTYPE)
    v["pos"] = pos
    if normal is not None:
        v["normal"] = normal
    if uv is not None:
        v["uv"] = uv
    return v


class SceneData:
    """Host scene in the reference's layouts."""

    def __init__(self):
        self.geoms = []        # (vertices, indices[n,3] u32, mat_index[n] u32, flags)
        self.materials = []
        self.instances = []    # (transform[4,4] column-major, geometry index)
        self.textures = []     # uint8 [h, w, 4]
        self.dir_light = abi.DirLight()
        self.point_lights = abi.PointLights()
        self.active_lights = abi.ActiveLights()
        self.env_kind = abi.KF_ENV_CLEAR_COLOR
        self.env_gradient = np.zeros(8, np.float32)
        self.env_faces = None
        self.env_offsets = None  # batched environments: instance offsets, n_env + 1
        self.separate_geometry = False  # True: no BVH shared between geometries with the same triangles
        self._n_point = 0
        self._n_active = 0

    def add_material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add_geometry(self, verts, indices, mat, opaque=True, hidden=False):
        """Geometry with one material for all triangles (Geometry::setMaterial)."""
        indices = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        if np.ndim(mat) == 0 and not isinstance(mat, np.ndarray):
            mi = np.full(len(indices), int(mat), np.uint32)
        elif isinstance(mat, np.ndarray) and mat.dtype == abi.MATERIAL_DTYPE:
            mi = np.full(len(indices), self.add_material(mat), np.uint32)
        else:
            mi = np.asarray(mat, np.uint32)
        flags = (abi.KF_GEOM_OPAQUE if opaque else 0) | (abi.KF_GEOM_HIDDEN if hidden else 0)
        self.geoms.append((verts, indices, mi, flags))
        return len(self.geoms) - 1

    def add_instance(self, geometry, transform=None):
        self.instances.append((mat_identity() if transform is None else np.asarray(transform, np.float32),
                               geometry))
        return len(self.instances) - 1

    def add_texture(self, rgba8):
        self.textures.append(np.ascontiguousarray(rgba8, np.uint8))
        return len(self.textures) - 1

    def set_directional_light(self, direction, color=(1, 1, 1), strength=0.0, softness=0.0):
        d = np.asarray(direction, np.float64)
        d = (d / np.linalg.norm(d)).astype(np.float32)  # scene.cpp:254
        self.dir_light.direction[:] = (*d, softness)
        self.dir_light.rgbs[:] = (*color, strength)

    def add_point_light(self, position, color=(1, 1, 1), strength=0.0, radius=1.0):
        if self._n_point >= abi.MAX_POINT_LIGHTS:
            return False  # scene.hpp:159-164 warns and drops
        i = self._n_point
        self.point_lights.posr[i][:] = (*position, radius)
        self.point_lights.rgbs[i][:] = (*color, strength)
        self._n_point += 1
        return True

    def add_active_light(self, view, color=(1, 1, 1), fov=0.0, strength=0.0, softness=0.0, tex_id=-1):
        if self._n_active >= abi.MAX_ACTIVE_LIGHTS:
            return False
        v = np.ascontiguousarray(np.asarray(view, np.float32).ravel())
        col = np.asarray(color, np.float32)
        check(lib().kf_active_light_fill(C.byref(self.active_lights), self._n_active, ptr(v), ptr(col), fov,
                                         strength, softness, tex_id))
        self._n_active += 1
        return True

    def set_gradient_environment(self, bottom, top):
        self.env_kind = abi.KF_ENV_GRADIENT
        self.env_gradient[:] = (*bottom, 0.0, *top, 0.0)

    def environment(self, e):
        """The scene of environment e alone: the same geometries, materials,
        textures, lights and sky with only e's instances."""
        import copy
        sub = copy.copy(self)
        lo, hi = int(self.env_offsets[e]), int(self.env_offsets[e + 1])
        sub.instances = list(self.instances[lo:hi])
        sub.env_offsets = None
        return sub

    def num_envs(self):
        return 1 if self.env_offsets is None else len(self.env_offsets) - 1

    # -------------------------------------------------------------- flatten
    def arrays(self):
        ng = len(self.geoms)
        voff = np.zeros(ng + 1, np.uint32)
        toff = np.zeros(ng + 1, np.uint32)
        for g, (v, ix, mi, fl) in enumerate(self.geoms):
            voff[g + 1] = voff[g] + len(v)
            toff[g + 1] = toff[g] + len(ix)
        verts = np.concatenate([g[0] for g in self.geoms]) if ng else np.zeros(0, abi.VERTEX_DTYPE)
        idx = np.concatenate([g[1] for g in self.geoms]) if ng else np.zeros((0, 3), np.uint32)
        mi = np.concatenate([g[2] for g in self.geoms]) if ng else np.zeros(0, np.uint32)
        flags = np.array([g[3] for g in self.geoms], np.uint32)
        inst = np.zeros(len(self.instances), abi.INSTANCE_DTYPE)
        for i, (t, g) in enumerate(self.instances):
            inst[i]["transform"] = np.asarray(t, np.float32).ravel()
            inst[i]["geometry_index"] = g
        mats = np.array(self.materials, abi.MATERIAL_DTYPE) if self.materials else np.zeros(0, abi.MATERIAL_DTYPE)
        return dict(voff=voff, toff=toff, flags=flags, verts=np.ascontiguousarray(verts),
                    idx=np.ascontiguousarray(idx), mi=np.ascontiguousarray(mi), inst=inst,
                    mats=np.ascontiguousarray(mats))

    def desc(self):
        """kf_scene_desc pointing into arrays kept alive on the returned object."""
        a = self.arrays()
        d = abi.SceneDesc()
        keep = [a]
        d.num_geometries = len(self.geoms)
        d.geom_vertex_offset = ptr(a["voff"])
        d.geom_tri_offset = ptr(a["toff"])
        d.geom_flags = ptr(a["flags"]) if len(self.geoms) else None
        d.vertices = ptr(a["verts"]) if len(a["verts"]) else None
        d.indices = ptr(a["idx"]) if len(a["idx"]) else None
        d.mat_index = ptr(a["mi"]) if len(a["mi"]) else None
        d.num_instances = len(a["inst"])
        d.instances = ptr(a["inst"]) if len(a["inst"]) else None
        d.num_materials = len(a["mats"])
        d.materials = ptr(a["mats"]) if len(a["mats"]) else None
        texs = (abi.TextureDesc * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            texs[i].width, texs[i].height, texs[i].rgba8 = t.shape[1], t.shape[0], ptr(t)
        keep.append(texs)
        d.num_textures = len(self.textures)
        d.textures = C.addressof(texs)
        d.dir_light = C.addressof(self.dir_light)
        d.point_lights = C.addressof(self.point_lights)
        d.active_lights = C.addressof(self.active_lights)
        d.env_kind = self.env_kind
        if self.env_faces is not None:
            faces = (abi.TextureDesc * 6)()
            for i, t in enumerate(self.env_faces):
                faces[i].width, faces[i].height, faces[i].rgba8 = t.shape[1], t.shape[0], ptr(t)
            keep.append(faces)
            d.env_faces = C.addressof(faces)
        for i in range(8):
            d.env_gradient[i] = float(self.env_gradient[i])
        if self.env_offsets is not None:
            eo = np.ascontiguousarray(self.env_offsets, np.uint32)
            keep.append(eo)
            d.num_envs = len(eo) - 1
            d.env_instance_offset = ptr(eo)
        d.separate_geometry = int(bool(self.separate_geometry))
        d._keep = keep
        return d

    def n_triangles(self):
        return sum(len(g[1]) for g in self.geoms)


def camera_ubo(position, front, up, width, height, fx=None, fy=None, cx=None, cy=None, skew=0.0, aperture=0.0,
               focal_length=5.0):
    """CameraUBO of Scene::uploadCameraBuffer (scene.cpp:236-250)."""
    fx = width * 0.5 if fx is None else fx   # camera.cpp:13-15 defaults
    fy = width * 0.5 if fy is None else fy
    cx = width * 0.5 if cx is None else cx
    cy = height * 0.5 if cy is None else cy
    u = abi.CameraUBO()
    p = np.asarray(position, np.float32)
    f = np.asarray(front, np.float32)
    up_ = np.asarray(up, np.float32)
    check(lib().kf_camera_ubo_compute(ptr(p), ptr(f), ptr(up_), int(width), int(height), fx, fy, cx, cy, skew,
                                      aperture, focal_length, C.byref(u)))
    return u


def push_constants(spp=32, path_depth=8, frame_count=0, clear_color=(0.0, 0.0, 0.0, 1.0), use_env_map=False,
                   russian_roulette=True, rr_min_bounces=4):
    """RtPushConstants (context.cpp:605-614) with Config defaults (config.hpp:257-269)."""
    pc = abi.PushConstants()
    pc.clear_color[:] = clear_color
    pc.frame_count = frame_count
    pc.spp = spp
    pc.path_depth = path_depth
    pc.use_env_map = int(use_env_map)
    pc.russian_roulette = int(russian_roulette)
    pc.rr_min_bounces = rr_min_bounces
    pc.nee = 1
    pc.nee_min_bounces = 0
    return pc


class GpuScene:
    """Scene uploaded to HBM with BLAS + TLAS built on the GPU."""

    def __init__(self, data):
        self.data = data
        self._desc = data.desc()
        h = C.c_void_p()
        check(lib().kf_scene_create(C.byref(self._desc), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().kf_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bvh_info(self):
        i = abi.BvhInfo()
        check(lib().kf_scene_bvh_info(self.h, C.byref(i)))
        return i

    def update_instances(self, inst):
        check(lib().kf_scene_update_instances(self.h, ptr(inst), len(inst)))

    def update_geometry(self, g, verts):
        """New vertices for geometry g (same count): refit its BLAS, rebuild the TLAS."""
        verts = np.ascontiguousarray(verts, abi.VERTEX_DTYPE)
        check(lib().kf_scene_update_geometry(self.h, g, ptr(verts), len(verts)))

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        hits = np.zeros(len(rays), abi.HIT_DTYPE)
        check(lib().kf_trace_closest(self.h, ptr(rays), len(rays), ptr(hits)))
        return hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        check(lib().kf_trace_any(self.h, ptr(rays), len(rays), ptr(occ)))
        return occ

    def read_blas(self, g):
        nn = C.c_uint32()
        root = C.c_int32()
        check(lib().kf_scene_read_blas(self.h, g, None, 0, C.byref(nn), C.byref(root), None, 0))
        nt = len(self.data.geoms[g][1])
        nodes = np.zeros((max(nn.value, 1), 16), np.float32)
        order = np.zeros(max(nt, 1), np.uint32)
        check(lib().kf_scene_read_blas(self.h, g, ptr(nodes), nodes.shape[0], C.byref(nn), C.byref(root),
                                       ptr(order), len(order)))
        return nodes[:nn.value], root.value, order[:nt]

    def read_bvh4(self, which):
        """BVH4 nodes (n, 32) float32 and root refs (per geometry for which=0,
        one per environment for which=1 = TLAS)."""
        nn = C.c_uint32()
        check(lib().kf_scene_read_bvh4(self.h, which, None, 0, C.byref(nn), None, 0))
        nodes = np.zeros((max(nn.value, 1), 32), np.float32)
        nr = max(len(self.data.geoms), 1) if which == 0 else self.data.num_envs()
        roots = np.zeros(nr, np.int32)
        check(lib().kf_scene_read_bvh4(self.h, which, ptr(nodes), nodes.shape[0], C.byref(nn), ptr(roots), nr))
        return nodes[:nn.value], roots


class Renderer:
    """RGBA32F render targets + one-frame render (RayTracer::trace)."""

    def __init__(self, width, height):
        self.w, self.h_ = width, height
        h = C.c_void_p()
        check(lib().kf_renderer_create(width, height, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().kf_renderer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, scene, cam, pc, seed=0, samples_per_batch=0, keep_samples=False, count_traversal=False,
               use_graph=False, full_tlas_camera=False):
        o = abi.RenderOpts(seed, samples_per_batch, int(keep_samples), int(count_traversal), int(use_graph),
                           int(full_tlas_camera))
        st = abi.RenderStats()
        check(lib().kf_render(self.h, scene.h, C.byref(cam), C.byref(pc), C.byref(o), C.byref(st)))
        return st

    def render_envs(self, scene, cams, pc, seed=0, samples_per_batch=0, keep_samples=False, count_traversal=False,
                    use_graph=False, full_tlas_camera=False):
        """One frame of every environment of a batched scene: cams[e] views
        environment e, whose image is rows [e*H/n, (e+1)*H/n) of the targets."""
        arr = (abi.CameraUBO * len(cams))(*cams)
        o = abi.RenderOpts(seed, samples_per_batch, int(keep_samples), int(count_traversal), int(use_graph),
                           int(full_tlas_camera))
        st = abi.RenderStats()
        check(lib().kf_render_envs(self.h, scene.h, len(cams), arr, C.byref(pc), C.byref(o), C.byref(st)))
        return st

    def autotune(self, scene, cams, pc):
        """Pick the wavefront schedule (samples per batch, refill threshold,
        reservation per wave) by timing frames; later renders with
        samples_per_batch=0 use it."""
        cams = cams if isinstance(cams, (list, tuple)) else [cams]
        arr = (abi.CameraUBO * len(cams))(*cams)
        t = abi.TuneResult()
        check(lib().kf_renderer_autotune(self.h, scene.h, len(cams), arr, C.byref(pc), C.byref(t)))
        return dict(samples_per_batch=t.samples_per_batch, refill=t.refill, grab=t.grab, tail_bounce=t.tail_bounce,
                    grab0=t.grab0, grab_shadow=t.grab_shadow, tail_pairs=t.tail_pairs, frame_ms=t.frame_ms,
                    default_ms=t.default_ms, trials=t.trials)

    def set_schedule(self, samples_per_batch=0, refill=0, grab=0, tail_bounce=0, grab0=0, grab_shadow=0,
                     tail_pairs=1):
        """grab0 / grab_shadow: reservation per wave for the camera / shadow rays (0: grab);
        tail_pairs 1: two lanes per path may carry the tail bounces, 0: one lane."""
        check(lib().kf_renderer_set_schedule(self.h, samples_per_batch, refill, grab, tail_bounce, grab0,
                                             grab_shadow))
        check(lib().kf_renderer_set_tail_pairs(self.h, tail_pairs))

    def read(self, target=0):
        out = np.zeros((self.h_, self.w, 4), np.float32)
        check(lib().kf_renderer_read(self.h, target, ptr(out)))
        return out

    def write(self, target, img):
        img = np.ascontiguousarray(img, np.float32)
        check(lib().kf_renderer_write(self.h, target, ptr(img)))

    def read_samples(self, spp):
        out = np.zeros((spp, self.h_, self.w, 3), np.float32)
        check(lib().kf_renderer_read_samples(self.h, ptr(out), out.size))
        return out

    def download_bgra8(self):
        out = np.zeros((self.h_, self.w, 4), np.uint8)
        check(lib().kf_renderer_download_bgra8(self.h, ptr(out)))
        return out

    def download_gray32f(self):
        """Single-channel FP32 image (r + g + b) / 3 of the rgba target."""
        out = np.zeros((self.h_, self.w), np.float32)
        check(lib().kf_renderer_download_gray32f(self.h, ptr(out)))
        return out

    def device_target(self, target):
        return lib().kf_renderer_device_target(self.h, target)
