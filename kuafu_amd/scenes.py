"""Workload scenes: meshes, the benchmark configurations and small test scenes.

Mesh generators follow the reference's conventions: outward winding
(cross(v1 - v0, v2 - v0) points out of the surface, as in createSphere /
createCube, geometry.cpp:297-537), smooth vertex normals.  Random scene
content uses numpy's PCG64 generator seeded explicitly.
"""
import math

import numpy as np

from . import abi
from .api import (SceneData, camera_ubo, look_at, mat_from_rows, mat_identity, mat_rotate, mat_scale,
                  mat_translate, material, push_constants, vertices)


# ------------------------------------------------------------------ meshes
def uv_sphere(slices=32, stacks=16, radius=1.0):
    """UV sphere, +Y poles: slices*2 + slices*(stacks-2)*2 triangles (960 for 32x16)."""
    pos = [(0.0, radius, 0.0)]
    uv = [(0.5, 0.0)]
    for i in range(1, stacks):
        phi = math.pi * i / stacks
        for j in range(slices):
            th = 2 * math.pi * j / slices
            pos.append((radius * math.sin(phi) * math.cos(th), radius * math.cos(phi), radius * math.sin(phi) * math.sin(th)))
            uv.append((j / slices, i / stacks))
    pos.append((0.0, -radius, 0.0))
    uv.append((0.5, 1.0))
    S = len(pos) - 1

    def ring(i, j):
        return 1 + (i - 1) * slices + (j % slices)

    tris = []
    for j in range(slices):
        tris.append((0, ring(1, j + 1), ring(1, j)))
    for i in range(1, stacks - 1):
        for j in range(slices):
            a, b = ring(i, j), ring(i, j + 1)
            c, d = ring(i + 1, j), ring(i + 1, j + 1)
            tris.append((a, b, d))
            tris.append((a, d, c))
    for j in range(slices):
        tris.append((S, ring(stacks - 1, j), ring(stacks - 1, j + 1)))
    pos = np.array(pos, np.float32)
    tris = np.array(tris, np.uint32)
    nrm = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    return _orient_outward(vertices(pos, nrm, np.array(uv, np.float32)), tris)


def icosphere(subdiv=4, radius=1.0):
    """Icosphere: 20 * 4^subdiv triangles (5120 for subdiv 4), 10 * 4^subdiv + 2 vertices."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    pos = (np.array(v) * radius).astype(np.float32)
    nrm = np.array(v, np.float32)
    uvs = np.stack([0.5 + np.arctan2(nrm[:, 2], nrm[:, 0]) / (2 * np.pi), 0.5 - np.arcsin(np.clip(nrm[:, 1], -1, 1)) / np.pi], 1)
    return _orient_outward(vertices(pos, nrm, uvs.astype(np.float32)), np.array(f, np.uint32))


def quad(size=1.0):
    """Square in the XY plane centred at the origin, normal +Z (2 triangles)."""
    h = size / 2
    pos = np.array([(-h, -h, 0), (h, -h, 0), (h, h, 0), (-h, h, 0)], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    uv = np.array([(0, 0), (1, 0), (1, 1), (0, 1)], np.float32)
    return vertices(pos, nrm, uv), np.array([(0, 1, 2), (0, 2, 3)], np.uint32)


def box(size=(1, 1, 1)):
    """Axis-aligned box centred at the origin with flat faces (12 triangles)."""
    sx, sy, sz = (s / 2 for s in size)
    pos, nrm, tris = [], [], []
    faces = [((1, 0, 0), (0, 1, 0), (0, 0, 1)), ((-1, 0, 0), (0, 0, 1), (0, 1, 0)), ((0, 1, 0), (0, 0, 1), (1, 0, 0)),
             ((0, -1, 0), (1, 0, 0), (0, 0, 1)), ((0, 0, 1), (1, 0, 0), (0, 1, 0)), ((0, 0, -1), (0, 1, 0), (1, 0, 0))]
    ext = np.array([sx, sy, sz])
    for n, u, w in faces:
        n, u, w = (np.array(a, np.float64) for a in (n, u, w))
        c = n * ext
        b = len(pos)
        for a, bb in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
            pos.append(c + a * u * ext + bb * w * ext)
            nrm.append(n)
        tris += [(b, b + 1, b + 2), (b, b + 2, b + 3)]
    return _orient_outward(vertices(np.array(pos, np.float32), np.array(nrm, np.float32)), np.array(tris, np.uint32))


def cylinder(segments=16, radius=0.5, height=1.0):
    """Capped cylinder along +Z centred at the origin: 4 * segments triangles."""
    pos, nrm, tris = [], [], []
    h = height / 2
    for j in range(segments):
        th = 2 * math.pi * j / segments
        c, s = math.cos(th), math.sin(th)
        pos += [(radius * c, radius * s, -h), (radius * c, radius * s, h)]
        nrm += [(c, s, 0), (c, s, 0)]
    for j in range(segments):
        a, b = 2 * j, 2 * ((j + 1) % segments)
        tris += [(a, b, b + 1), (a, b + 1, a + 1)]
    top = len(pos)
    pos.append((0, 0, h)); nrm.append((0, 0, 1))
    bot = len(pos)
    pos.append((0, 0, -h)); nrm.append((0, 0, -1))
    cap0 = len(pos)
    for j in range(segments):
        th = 2 * math.pi * j / segments
        pos.append((radius * math.cos(th), radius * math.sin(th), h)); nrm.append((0, 0, 1))
    cap1 = len(pos)
    for j in range(segments):
        th = 2 * math.pi * j / segments
        pos.append((radius * math.cos(th), radius * math.sin(th), -h)); nrm.append((0, 0, -1))
    for j in range(segments):
        tris.append((top, cap0 + j, cap0 + (j + 1) % segments))
        tris.append((bot, cap1 + (j + 1) % segments, cap1 + j))
    return _orient_outward(vertices(np.array(pos, np.float32), np.array(nrm, np.float32)), np.array(tris, np.uint32))


def _orient_outward(v, tris):
    """Flip triangles whose geometric normal disagrees with the vertex normals."""
    p = v["pos"].astype(np.float64)
    n = v["normal"].astype(np.float64)
    g = np.cross(p[tris[:, 1]] - p[tris[:, 0]], p[tris[:, 2]] - p[tris[:, 0]])
    s = (g * (n[tris[:, 0]] + n[tris[:, 1]] + n[tris[:, 2]])).sum(1)
    tris = tris.copy()
    flip = s < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return v, np.ascontiguousarray(tris, np.uint32)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ------------------------------------------------------------------ workloads
class Workload:
    def __init__(self, name, data, cam, pc, width, height, seed=0):
        self.name, self.data, self.cam, self.pc = name, data, cam, pc
        self.width, self.height, self.seed = width, height, seed


def config1(width=256, height=256, spp=4, path_depth=4, seed=0):
    """Plain tracer, 256x256, 4 spp, depth 4: ground plane + 5 UV spheres (32x16)."""
    s = SceneData()
    ground_v, ground_i = quad(10.0)
    # quad lies in XY with normal +Z; rotate to the XZ plane with normal +Y
    g = s.add_geometry(ground_v, ground_i, material(diffuse=(0.8, 0.8, 0.8)))
    s.add_instance(g, mat_rotate(-math.pi / 2, (1, 0, 0)))
    sv, si = uv_sphere(32, 16, 0.5)
    mats = [material(diffuse=(0.8, 0.2, 0.2)),
            material(specular=1.0, roughness=0.05),
            material(metallic=1.0, roughness=0.5),
            material(transmission=1.0, ior=1.45),
            material(emission=(1.0, 1.0, 1.0), emission_strength=2.0)]
    for k, m in enumerate(mats):
        gi = s.add_geometry(sv, si, m)
        s.add_instance(gi, mat_translate(((k - 2) * 1.2, 0.5, 0.0)))
    s.set_directional_light((-0.3, -1.0, -0.2), (1, 1, 1), 3.0, 0.0)
    s.add_point_light((0.0, 3.0, 0.0), (1, 1, 1), 10.0, radius=0.0)
    eye = np.array([0.0, 2.0, 6.0])
    front = np.array([0.0, 0.5, 0.0]) - eye
    front /= np.linalg.norm(front)
    cam = camera_ubo(eye, front, (0, 1, 0), width, height, fx=width * 1.0, fy=width * 1.0)
    pc = push_constants(spp=spp, path_depth=path_depth, clear_color=(1.0, 1.0, 1.0, 0.1))
    return Workload("config1_spheres_256x256_4spp", s, cam, pc, width, height, seed)


def config2(width=1920, height=1080, spp=16, path_depth=8, seed=0, n_inst=256, subdiv=4):
    """Plain tracer, 1080p, 16 spp, depth 8: 256 icosphere instances (subdiv 4), area + spot light,
    gradient environment."""
    rng = np.random.default_rng(seed)
    s = SceneData()
    iv, ii = icosphere(subdiv)
    for k in range(n_inst):
        albedo = rng.uniform(0, 1, 3)
        rough = rng.uniform(0.02, 1.0)
        metal = 1.0 if rng.uniform() < 0.3 else 0.0
        trans = 1.0 if rng.uniform() < 0.15 else 0.0
        ior = rng.uniform(1.3, 1.7)
        m = material(diffuse=tuple(albedo), roughness=rough, metallic=metal, transmission=trans, ior=ior)
        gi = s.add_geometry(iv, ii, m)
        R = random_rotation(rng) * rng.uniform(0.2, 1.5)
        t = rng.uniform(-10.0, 10.0, 3)
        s.add_instance(gi, mat_from_rows(R, t))
    # 1 x 1 m area light (emissive quad) above the cube, facing down (-Z)
    qv, qi = quad(1.0)
    lg = s.add_geometry(qv, qi, material(emission=(1.0, 1.0, 1.0), emission_strength=50.0))
    s.add_instance(lg, mat_rotate(math.pi, (1, 0, 0), mat_translate((0.0, 0.0, 11.0))))
    # spot light: active light without pattern, 30 degree cone (scene.cpp:407-411)
    view = look_at((0.0, -16.0, 8.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    s.add_active_light(view, (1.0, 0.95, 0.9), fov=math.radians(30.0), strength=600.0)
    s.set_gradient_environment((0.05, 0.05, 0.06), (0.35, 0.45, 0.65))
    eye = np.array([0.0, -30.0, 6.0])
    front = -eye / np.linalg.norm(eye)
    cam = camera_ubo(eye, front, (0, 0, 1), width, height, fx=width * 0.62, fy=width * 0.62)
    pc = push_constants(spp=spp, path_depth=path_depth, use_env_map=True)
    return Workload("config2_icospheres256_1080p_16spp", s, cam, pc, width, height, seed)


def config3(width=1280, height=720, spp=32, path_depth=6, seed=0, camera="left"):
    """Active light sensing, 720p: 64 boxes/cylinders on a plane, IR projector with a
    Bernoulli(0.1) dot pattern; left/right IR cameras 55 mm apart and an RGB camera."""
    rng = np.random.default_rng(seed)
    s = SceneData()
    pv, pi = quad(4.0)
    g = s.add_geometry(pv, pi, material(diffuse=(0.7, 0.7, 0.7), roughness=0.8))
    s.add_instance(g)
    for k in range(64):
        size = rng.uniform(0.05, 0.4, 3)
        m = material(diffuse=tuple(rng.uniform(0.2, 0.9, 3)), roughness=rng.uniform(0.1, 1.0),
                     metallic=1.0 if rng.uniform() < 0.2 else 0.0)
        if rng.uniform() < 0.5:
            v, i = box(size)
        else:
            seg = int(rng.integers(3, 33))
            v, i = cylinder(seg, size[0] / 2, size[2])
        gi = s.add_geometry(v, i, m)
        pos = (rng.uniform(-1.2, 1.2), rng.uniform(-1.2, 1.2), size[2] / 2)
        s.add_instance(gi, mat_rotate(rng.uniform(0, 2 * math.pi), (0, 0, 1), mat_translate(pos)))
    pattern = (rng.uniform(size=(height, width)) < 0.1).astype(np.uint8) * 255
    tex = np.stack([pattern, pattern, pattern, np.full_like(pattern, 255)], -1)
    tid = s.add_texture(tex)
    proj_pos = np.array([0.0, -1.8, 1.8])
    view = look_at(proj_pos, (0.0, 0.0, 0.0), (0, 0, 1))
    s.add_active_light(view, (1.0, 1.0, 1.0), fov=math.radians(70.0), strength=8.0, tex_id=tid)
    s.set_directional_light((0.2, 0.3, -1.0), (1, 1, 1), 0.5, 0.0)
    offs = {"left": -0.0275, "right": 0.0275, "rgb": 0.0}
    eye = proj_pos + np.array([offs[camera], 0.0, 0.0])
    front = -proj_pos / np.linalg.norm(proj_pos)
    cam = camera_ubo(eye, front, (0, 0, 1), width, height, fx=width * 0.7, fy=width * 0.7)
    pc = push_constants(spp=spp, path_depth=path_depth, clear_color=(0.0, 0.0, 0.0, 1.0))
    return Workload("config3_active_light_720p_32spp_" + camera, s, cam, pc, width, height, seed)


def config5(n_env=128, size=512, spp=8, path_depth=6, seed=0, tri_range=(5000, 400000), n_assets=160):
    """Batched environments: n_env SAPIEN-style rooms, one size x size camera each, 8 spp, depth 6.

    A pool of n_assets geometries (icospheres of subdivision 1-5, boxes, cylinders; materials are
    per geometry in the reference API, drawn per asset as in config2) is instanced into every
    environment on a 10 m ground plane until the environment's triangle budget, log-uniform in
    tri_range, is used up (larger assets are drawn in proportion to their size, so big budgets
    take tens of instances, not thousands).  Each environment has its own TLAS and camera; the sun,
    a point light and the gradient sky are shared.  The render target stacks the environments'
    images vertically (width size, height size * n_env)."""
    rng = np.random.default_rng(seed)
    s = SceneData()
    spheres = {k: icosphere(k) for k in (1, 2, 3, 4, 5)}  # 80 .. 20480 triangles
    asset_tris = []
    for a in range(n_assets):
        kind = rng.uniform()
        albedo = rng.uniform(0, 1, 3)
        m = material(diffuse=tuple(albedo), roughness=rng.uniform(0.02, 1.0),
                     metallic=1.0 if rng.uniform() < 0.3 else 0.0,
                     transmission=1.0 if rng.uniform() < 0.15 else 0.0, ior=rng.uniform(1.3, 1.7))
        if kind < 0.6:
            v, t = spheres[int(rng.integers(1, 6))]
        elif kind < 0.8:
            v, t = box(tuple(rng.uniform(0.5, 1.5, 3)))
        else:
            v, t = cylinder(int(rng.integers(6, 33)), 0.5, rng.uniform(0.5, 2.0))
        s.add_geometry(v, t, m)
        asset_tris.append(len(t))
    asset_tris = np.array(asset_tris)
    gv, gt = quad(10.0)
    grounds = [s.add_geometry(gv, gt, material(diffuse=tuple(rng.uniform(0.3, 0.8, 3)), roughness=0.9))
               for _ in range(8)]
    offsets = [0]
    cams = []
    env_tris = []
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    for e in range(n_env):
        budget = int(math.exp(rng.uniform(lo, hi)))
        s.add_instance(grounds[int(rng.integers(0, len(grounds)))])
        used = 2
        while True:
            fit = np.nonzero(asset_tris <= budget - used)[0]
            if len(fit) == 0:
                break
            w = asset_tris[fit].astype(np.float64)
            a = int(fit[rng.choice(len(fit), p=w / w.sum())])
            scale = rng.uniform(0.15, 0.5)
            R = random_rotation(rng) * scale
            t = (rng.uniform(-2.0, 2.0), rng.uniform(-2.0, 2.0), scale)
            s.add_instance(a, mat_from_rows(R, t))
            used += int(asset_tris[a])
        offsets.append(len(s.instances))
        env_tris.append(used)
        az, el, dist = rng.uniform(0, 2 * math.pi), rng.uniform(0.35, 0.7), rng.uniform(5.0, 6.5)
        eye = np.array([dist * math.cos(el) * math.cos(az), dist * math.cos(el) * math.sin(az), dist * math.sin(el)])
        target = np.array([0.0, 0.0, 0.3])
        front = (target - eye) / np.linalg.norm(target - eye)
        cams.append(camera_ubo(eye, front, (0, 0, 1), size, size, fx=size * 0.9, fy=size * 0.9))
    s.env_offsets = np.array(offsets, np.uint32)
    s.set_directional_light((-0.3, -0.2, -1.0), (1.0, 0.97, 0.9), 2.5, 0.0)
    s.add_point_light((0.0, 0.0, 4.0), (1.0, 1.0, 1.0), 20.0, radius=0.0)
    s.set_gradient_environment((0.05, 0.05, 0.06), (0.35, 0.45, 0.65))
    pc = push_constants(spp=spp, path_depth=path_depth, use_env_map=True)
    w = Workload("config5_envs%d_%dx%d_%dspp" % (n_env, size, size, spp), s, cams, pc, size, size * n_env, seed)
    w.n_env, w.env_height, w.env_tris = n_env, size, np.array(env_tris)
    return w


def tiny_scene(seed=0, n_spheres=3, with_lights=True, soft=False):
    """Small randomized scene for parity tests (a few hundred triangles)."""
    rng = np.random.default_rng(seed)
    s = SceneData()
    gv, gi = quad(8.0)
    g = s.add_geometry(gv, gi, material(diffuse=(0.7, 0.7, 0.7), roughness=0.6))
    s.add_instance(g, mat_rotate(-math.pi / 2, (1, 0, 0)))
    sv, si = uv_sphere(12, 8, 0.5)
    kinds = [dict(diffuse=(0.8, 0.3, 0.2)), dict(specular=1.0, roughness=0.1), dict(metallic=1.0, roughness=0.4),
             dict(transmission=1.0, ior=1.5), dict(emission=(1, 1, 1), emission_strength=3.0),
             dict(diffuse=(0.2, 0.6, 0.9), metallic=0.5, roughness=0.3)]
    for k in range(n_spheres):
        kw = kinds[(k + seed) % len(kinds)]
        m = s.add_geometry(sv, si, material(**kw))
        R = random_rotation(rng) * rng.uniform(0.6, 1.4)
        s.add_instance(m, mat_from_rows(R, (rng.uniform(-2, 2), rng.uniform(0.4, 1.2), rng.uniform(-2, 2))))
    if with_lights:
        s.set_directional_light((-0.3, -1.0, -0.4), (1, 0.9, 0.8), 2.0, 0.3 if soft else 0.0)
        s.add_point_light((0.5, 2.5, 0.5), (1, 1, 1), 6.0, radius=0.3 if soft else 0.0)
        view = look_at((2.0, 3.0, 2.0), (0.0, 0.0, 0.0), (0, 1, 0))
        s.add_active_light(view, (0.9, 0.9, 1.0), fov=math.radians(50), strength=20.0, softness=0.2 if soft else 0.0)
    return s


def training_scene(seed, width=256, height=256):
    """Randomised scene for denoiser training: the config-1 generator (ground
    plane, UV spheres of the five material kinds, directional + point light,
    clear colour) with random counts, materials, lights and camera; one in four
    is a config-2-style cluster of icospheres under a gradient environment."""
    rng = np.random.default_rng(1000003 * seed + 17)
    s = SceneData()
    kinds = ["diffuse", "specular", "metal", "glass", "emissive", "mixed"]

    def rand_material(kind):
        col = tuple(rng.uniform(0.05, 0.95, 3))
        if kind == "diffuse":
            return material(diffuse=col, roughness=rng.uniform(0.3, 1.0))
        if kind == "specular":
            return material(diffuse=col, specular=rng.uniform(0.5, 1.0), roughness=rng.uniform(0.02, 0.3))
        if kind == "metal":
            return material(diffuse=col, metallic=1.0, roughness=rng.uniform(0.05, 0.8))
        if kind == "glass":
            return material(diffuse=tuple(rng.uniform(0.7, 1.0, 3)), transmission=1.0, ior=rng.uniform(1.3, 1.7))
        if kind == "emissive":
            return material(emission=tuple(rng.uniform(0.5, 1.0, 3)), emission_strength=rng.uniform(1.0, 6.0))
        return material(diffuse=col, metallic=rng.uniform(0, 1), roughness=rng.uniform(0.05, 1.0),
                        transmission=float(rng.uniform() < 0.2), ior=rng.uniform(1.3, 1.7))

    if rng.uniform() < 0.75:
        gv, gi = quad(rng.uniform(6.0, 14.0))
        g = s.add_geometry(gv, gi, material(diffuse=tuple(rng.uniform(0.2, 0.9, 3)), roughness=rng.uniform(0.2, 1.0),
                                            metallic=float(rng.uniform() < 0.1)))
        s.add_instance(g, mat_rotate(-math.pi / 2, (1, 0, 0)))
        sv, si = uv_sphere(32, 16, 0.5)
        for k in range(int(rng.integers(3, 9))):
            m = s.add_geometry(sv, si, rand_material(kinds[int(rng.integers(0, len(kinds)))]))
            r = rng.uniform(0.4, 1.6)
            s.add_instance(m, mat_scale(r, mat_translate((rng.uniform(-2.5, 2.5), 0.5 * r, rng.uniform(-2.5, 1.5)))))
        s.set_directional_light(tuple(rng.uniform(-1, 1, 3) * (1, 0, 1) + (0, -1, 0)), tuple(rng.uniform(0.7, 1, 3)),
                                rng.uniform(0.5, 4.0), 0.0)
        if rng.uniform() < 0.7:
            s.add_point_light((rng.uniform(-2, 2), rng.uniform(1.5, 4), rng.uniform(-2, 2)), tuple(rng.uniform(0.7, 1, 3)),
                              rng.uniform(2, 15), radius=0.0)
        eye = np.array([rng.uniform(-3, 3), rng.uniform(0.8, 3.5), rng.uniform(4, 8)])
        tgt = np.array([rng.uniform(-0.5, 0.5), rng.uniform(0.2, 0.8), 0.0])
        up = (0, 1, 0)
        cc = tuple(rng.uniform(0.3, 1.0, 3)) + (rng.uniform(0.02, 0.6),)
        pc = push_constants(spp=1, path_depth=int(rng.integers(4, 9)), clear_color=cc)
    else:
        iv, ii = icosphere(3)
        for k in range(int(rng.integers(12, 40))):
            m = s.add_geometry(iv, ii, rand_material(kinds[int(rng.integers(0, len(kinds)))]))
            R = random_rotation(rng) * rng.uniform(0.2, 1.2)
            s.add_instance(m, mat_from_rows(R, rng.uniform(-4, 4, 3)))
        view = look_at(tuple(rng.uniform(-8, 8, 3)), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
        s.add_active_light(view, (1.0, 0.95, 0.9), fov=math.radians(rng.uniform(20, 60)), strength=rng.uniform(50, 400))
        s.set_gradient_environment(tuple(rng.uniform(0.0, 0.1, 3)), tuple(rng.uniform(0.2, 0.8, 3)))
        d = rng.normal(size=3)
        eye = 12.0 * d / np.linalg.norm(d)
        tgt = np.zeros(3)
        up = (0, 0, 1)
        pc = push_constants(spp=1, path_depth=8, use_env_map=True)
    front = (tgt - eye) / np.linalg.norm(tgt - eye)
    f = width * rng.uniform(0.7, 1.4)
    cam = camera_ubo(eye, front, up, width, height, fx=f, fy=f)
    return Workload("train%d" % seed, s, cam, pc, width, height, seed)
