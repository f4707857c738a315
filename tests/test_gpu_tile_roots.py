"""Camera rays from per-tile TLAS roots (marked gpu).

k_tile_roots gives each 8x8 pixel tile a root item over just the instances
whose world boxes meet the tile's pyramid of camera rays.  The closest hit
does not depend on the order in which the traversal meets records (smallest
t, ties by instance and primitive), so every frame must equal, bit for bit,
the frame whose camera rays start from the whole TLAS
(kf_render_opts.full_tlas_camera): per-sample radiance, the three targets
and the ray counts, with fewer TLAS node visits where tiles cull instances.

Covered: the config-2 icosphere field (tiles of 0, 1, 2-4, 5-16 and more
candidates), a raster whose sides are not multiples of 8, batched
environments (an empty one, a one-instance one, several-instance ones),
a thin-lens camera (every tile keeps its environment's root), alpha-tested
and soft-shadow scenes, a camera inside an instance's box, and frame graph
replays.
"""
import numpy as np
import pytest

from kuafu_amd import scenes
from kuafu_amd.api import GpuScene, Renderer, camera_ubo, push_constants
from test_gpu_envs import _tail_workloads, same, small_envs

pytestmark = pytest.mark.gpu


def _frame(gs, w, full, seed=3, graph=False):
    rd = Renderer(w.width, w.height)
    pc = w.pc
    pc.frame_count = 0
    st = rd.render(gs, w.cam, pc, seed=seed, keep_samples=not graph, count_traversal=not graph, use_graph=graph,
                   full_tlas_camera=full)
    smp = None if graph else rd.read_samples(pc.spp)
    return st, smp, [rd.read(k) for k in range(3)]


def _assert_same_frames(w, expect_fewer=True):
    gs = GpuScene(w.data)
    st0, smp0, tg0 = _frame(gs, w, True)
    st1, smp1, tg1 = _frame(gs, w, False)
    print("%s: TLAS node visits %d (whole TLAS) -> %d (tile roots); all nodes %d -> %d"
          % (w.name, st0.tlas_node_visits, st1.tlas_node_visits, st0.node_visits, st1.node_visits), flush=True)
    assert same(smp1, smp0), w.name
    for k in range(3):
        assert same(tg1[k], tg0[k]), "%s target %d" % (w.name, k)
    assert (st1.primary_rays, st1.extension_rays, st1.shadow_rays) == \
        (st0.primary_rays, st0.extension_rays, st0.shadow_rays)
    if expect_fewer:
        assert st1.tlas_node_visits < st0.tlas_node_visits
    return st0, st1


def test_tile_roots_config2_field(gpu):
    # 256 icosphere instances: tiles see none, one, a few or many of them
    w = scenes.config2(width=240, height=136, spp=2, path_depth=4)
    st0, st1 = _assert_same_frames(w)
    assert st1.node_visits < st0.node_visits


def test_tile_roots_ragged_raster(gpu):
    w = scenes.config2(width=101, height=59, spp=2, path_depth=4, n_inst=60, subdiv=2)
    _assert_same_frames(w)


def test_tile_roots_small_scenes(gpu):
    # config-1 spheres, a small config-2 field, inline soft shadows, alpha-tested spheres
    for w in _tail_workloads():
        _assert_same_frames(w, expect_fewer=False)


def test_tile_roots_thin_lens_camera(gpu):
    w = scenes.config2(width=64, height=40, spp=2, path_depth=3, n_inst=40, subdiv=2)
    eye = np.array(w.cam.position[:3], np.float32)
    front = np.array(w.cam.front[:3], np.float32)
    w.cam = camera_ubo(eye, front, (0, 0, 1), w.width, w.height, fx=w.width * 0.62, fy=w.width * 0.62,
                       aperture=0.2, focal_length=8.0)
    assert w.cam.position[3] > 0
    st0, st1 = _assert_same_frames(w, expect_fewer=False)
    assert st1.tlas_node_visits == st0.tlas_node_visits  # rays from a disc: every tile keeps the whole TLAS


def test_tile_roots_camera_inside_instance_box(gpu):
    # a large sphere around the camera (its box holds the camera position)
    # and small spheres in front: the enclosing instance is in every tile
    data = scenes.tiny_scene(4, n_spheres=4)
    sv, si = scenes.uv_sphere(24, 12, 12.0)
    g = data.add_geometry(sv, si, scenes.material(diffuse=(0.5, 0.5, 0.6)))
    data.add_instance(g, scenes.mat_from_rows(np.eye(3), (0.0, 1.0, 2.0)))
    cam = camera_ubo((0.0, 2.0, 5.0), (0, -0.3, -1.0), (0, 1, 0), 56, 40, fx=40.0, fy=40.0)
    w = scenes.Workload("inside", data, cam, push_constants(spp=2, path_depth=4), 56, 40)
    _assert_same_frames(w, expect_fewer=False)


def test_tile_roots_batched_environments(gpu):
    w, d, cams = small_envs()
    gs = GpuScene(d)
    n = len(cams)
    out = []
    for full in (True, False):
        rd = Renderer(w.width, w.height * n)
        pc = w.pc
        pc.frame_count = 0
        st = rd.render_envs(gs, cams, pc, seed=4, count_traversal=True, full_tlas_camera=full)
        out.append((st, [rd.read(k) for k in range(3)]))
    (st0, tg0), (st1, tg1) = out
    print("environments: TLAS node visits %d -> %d" % (st0.tlas_node_visits, st1.tlas_node_visits))
    for k in range(3):
        assert same(tg1[k], tg0[k]), "target %d" % k
    assert (st1.primary_rays, st1.extension_rays, st1.shadow_rays) == \
        (st0.primary_rays, st0.extension_rays, st0.shadow_rays)
    assert st1.tlas_node_visits < st0.tlas_node_visits


def test_tile_roots_graph_replays(gpu):
    # captured frames (the tile roots are rebuilt inside the graph from the
    # cameras of each replay) against eager frames, after a camera move
    w = scenes.config2(width=96, height=56, spp=2, path_depth=4, n_inst=60, subdiv=2)
    gs = GpuScene(w.data)
    rg = Renderer(w.width, w.height)
    re_ = Renderer(w.width, w.height)
    base = w.cam
    eye = np.array(base.position[:3], np.float32)
    front = np.array(base.front[:3], np.float32)
    moved = camera_ubo(eye + np.float32([1.5, 0.3, -2.0]), front, (0, 0, 1), w.width, w.height)
    pc = w.pc
    for f, cam in enumerate((base, moved, base)):
        pc.frame_count = 0
        rg.render(gs, cam, pc, seed=f, use_graph=True)
        re_.render(gs, cam, pc, seed=f, full_tlas_camera=True)
        for k in range(3):
            assert same(rg.read(k), re_.read(k)), "frame %d target %d" % (f, k)
    assert rg.render(gs, base, pc, seed=9, use_graph=True).graph_captures == 1
