"""Batched environments on the GPU (marked gpu).

A scene of several environments, each with its own TLAS tree and camera, is
rendered in one frame by kf_render_envs.  Environment e's rows of the stacked
target must equal, bit for bit, a plain kf_render of a scene holding only
e's instances (same seeds, same W x He raster), and match the FP32 oracle on
that scene.  Covered root kinds: an inner BVH4 node (several instances), a
single instance leaf, an empty environment.
"""
import numpy as np
import pytest

import oracle
from kuafu_amd import abi, scenes
from kuafu_amd.api import GpuScene, Renderer

pytestmark = pytest.mark.gpu


def small_envs():
    """Three generated environments plus an empty one and a one-instance one:
    instance ranges [0,a) [a,a) [a,b) [b,c) [c,c+1)."""
    w = scenes.config5(n_env=3, size=32, spp=2, path_depth=3, seed=5, tri_range=(300, 6000), n_assets=12)
    d = w.data
    a, b, c = (int(x) for x in d.env_offsets[1:])
    d.instances.append(d.instances[0])  # a ground plane alone
    d.env_offsets = np.array([0, a, a, b, c, c + 1], np.uint32)
    cams = [w.cam[0], w.cam[1], w.cam[1], w.cam[2], w.cam[0]]
    return w, d, cams


def same(x, y):
    return ((x == y) | (np.isnan(x) & np.isnan(y))).all()


def test_env_bvh4_trees(gpu):
    """Every environment's TLAS BVH4 holds exactly its own instances, its
    nodes inside its own range of the array, each reached once."""
    from test_gpu_parity import walk_bvh4
    w, d, _ = small_envs()
    gs = GpuScene(d)
    nodes, roots = gs.read_bvh4(1)
    off = [int(x) for x in d.env_offsets]
    node_off = np.cumsum([0] + [max(off[e + 1] - off[e] - 1, 0) for e in range(len(off) - 1)])
    info = gs.bvh_info()
    assert info.tlas_nodes == node_off[-1] and info.instances == len(d.instances)
    total = 0
    for e in range(len(off) - 1):
        n_e = off[e + 1] - off[e]
        if n_e == 0:
            assert roots[e] == 0x7fffffff
            continue
        if n_e == 1:
            assert roots[e] == ~off[e]
            continue
        assert roots[e] == node_off[e]
        seen, leaves = walk_bvh4(nodes, int(roots[e]), lambda i, lo, hi: None, 0)
        used = np.nonzero(seen)[0]
        assert (seen[used] == 1).all()
        assert used.min() >= node_off[e] and used.max() < node_off[e + 1]
        assert sorted(leaves) == list(range(off[e], off[e + 1]))
        total += len(used)
    assert info.tlas4_nodes == total


def test_render_envs_equal_single_environment_renders(gpu):
    w, d, cams = small_envs()
    n_env, size = len(cams), w.env_height
    pc = w.pc
    pc.frame_count = 0
    gs = GpuScene(d)
    rd = Renderer(size, size * n_env)
    st = rd.render_envs(gs, cams, pc, seed=3, keep_samples=True)
    smp = rd.read_samples(pc.spp)
    tgt = [rd.read(k) for k in range(3)]
    sums = np.zeros(3, np.int64)
    for e in range(n_env):
        sub = d.environment(e)
        r1 = Renderer(size, size)
        st1 = r1.render(GpuScene(sub), cams[e], pc, seed=3, keep_samples=True)
        rows = slice(e * size, (e + 1) * size)
        assert same(smp[:, rows], r1.read_samples(pc.spp)), "environment %d samples" % e
        for k in range(3):
            assert same(tgt[k][rows], r1.read(k)), "environment %d target %d" % (e, k)
        sums += (st1.primary_rays, st1.extension_rays, st1.shadow_rays)
    assert (st.primary_rays, st.extension_rays, st.shadow_rays) == tuple(int(x) for x in sums)
    # the environments differ (distinct cameras and content)
    assert not same(tgt[0][0:size], tgt[0][2 * size:3 * size])
    # the empty environment shows the sky only; environment 3 against the oracle
    o = oracle.Scene(d.environment(3))
    ref = o.render(cams[3], pc, size, size, seed=3, threads=8, keep_samples=True)
    rs = ref["samples"]
    got = smp[:, 3 * size:4 * size]
    close = (np.abs(got - rs) <= 1e-4 * np.maximum(1.0, np.abs(rs))) | (np.isnan(got) & np.isnan(rs))
    assert close.all(-1).mean() >= 0.999


def test_render_envs_rejects_bad_calls(gpu):
    w, d, cams = small_envs()
    gs = GpuScene(d)
    n_env, size = len(cams), w.env_height
    rd = Renderer(size, size * n_env)
    with pytest.raises(RuntimeError, match="environments"):
        rd.render(gs, cams[0], w.pc)
    with pytest.raises(RuntimeError, match="number of cameras"):
        rd.render_envs(gs, cams[:2], w.pc)
    with pytest.raises(RuntimeError, match="multiple"):
        Renderer(size, size * n_env + 1).render_envs(gs, cams, w.pc)
    inst = d.arrays()["inst"]
    with pytest.raises(RuntimeError, match="instance count"):
        gs.update_instances(inst[:-1])
    # a batched scene keeps working after the rejected calls
    rd.render_envs(gs, cams, w.pc)
    assert np.isfinite(rd.read(0)).all()


def test_env_instance_update(gpu):
    """kf_scene_update_instances on a batched scene (every environment's
    tree rebuilt in one launch) equals a scene created with the new
    transforms."""
    w, d, cams = small_envs()
    gs = GpuScene(d)
    n_env, size = len(cams), w.env_height
    rng = np.random.default_rng(9)
    import copy
    moved = copy.copy(d)
    moved.instances = [(np.asarray(t, np.float32).copy(), g) for t, g in d.instances]
    for t, g in moved.instances[1:]:
        t[3, :2] += rng.uniform(-0.5, 0.5, 2).astype(np.float32)
    gs.update_instances(moved.arrays()["inst"])
    rd = Renderer(size, size * n_env)
    rd.render_envs(gs, cams, w.pc, seed=4)
    rd2 = Renderer(size, size * n_env)
    rd2.render_envs(GpuScene(moved), cams, w.pc, seed=4)
    assert same(rd.read(0), rd2.read(0))


def test_frame_graph_replays_match_eager_frames(gpu):
    """A captured frame graph (three batches of one sample, two streams)
    replayed over frames with new seeds, a moved camera and the running
    mean, interleaved with eager work, equals eager rendering bit for bit;
    a changed sample count or reservation schedule captures a new graph."""
    w = scenes.config2(width=48, height=27, spp=3, path_depth=4, n_inst=24, subdiv=2)
    gs = GpuScene(w.data)
    rg, re = Renderer(w.width, w.height), Renderer(w.width, w.height)
    from kuafu_amd.api import camera_ubo
    cam2 = camera_ubo((2.0, -28.0, 7.0), (-0.05, 1.0, -0.2), (0, 0, 1), w.width, w.height, fx=30.0, fy=30.0)
    probe = np.zeros(4, abi.RAY_DTYPE)
    probe["dir"] = (0.0, 1.0, 0.0)
    probe["org"] = (0.0, -30.0, 0.0)
    probe["tmax"] = 100.0
    pc = w.pc
    for f in range(7):
        pc.frame_count = f
        cam = cam2 if f in (2, 3) else w.cam
        if f == 5:
            pc.spp = 2
        if f == 6:  # camera- and shadow-ray reservations of their own
            rg.set_schedule(0, 0, 0, 0, 1024, 32)
        sg = rg.render(gs, cam, pc, seed=10 + f, samples_per_batch=1, use_graph=True)
        se = re.render(gs, cam, pc, seed=10 + f, samples_per_batch=1)
        gs.trace_closest(probe)  # eager work between replays
        assert same(rg.read(0), re.read(0)), "frame %d" % f
        assert (sg.primary_rays, sg.extension_rays, sg.shadow_rays) == (se.primary_rays, se.extension_rays,
                                                                        se.shadow_rays)
        assert sg.extend_launches == se.extend_launches and sg.trace_ms > 0
        # captured at frame 0, again when the sample count changes (frame 5)
        # and when the reservations change (frame 6); replayed otherwise
        assert sg.graph_captures == 1 + (f >= 5) + (f >= 6), (f, sg.graph_captures)
        assert se.graph_captures == 0


def test_frame_graph_batched_environments(gpu):
    w, d, cams = small_envs()
    gs = GpuScene(d)
    n_env, size = len(cams), w.env_height
    rg, re = Renderer(size, size * n_env), Renderer(size, size * n_env)
    pc = w.pc
    for f in range(3):
        pc.frame_count = f
        cs = cams[::-1] if f == 1 else cams  # other cameras, same graph
        rg.render_envs(gs, cs, pc, seed=f, use_graph=True)
        re.render_envs(gs, cs, pc, seed=f)
        assert same(rg.read(0), re.read(0)), "frame %d" % f


def test_autotune_keeps_images(gpu):
    w = scenes.config2(width=64, height=36, spp=4, path_depth=4, n_inst=24, subdiv=2)
    gs = GpuScene(w.data)
    rd = Renderer(w.width, w.height)
    pc = w.pc
    pc.frame_count = 0
    rd.render(gs, w.cam, pc, seed=2)
    before = rd.read(0).copy()
    t = rd.autotune(gs, w.cam, pc)
    print(t)
    assert t["samples_per_batch"] in (1, 2, 4) and t["refill"] in (8, 16, 32, 48)
    assert t["grab"] in (64, 128, 256, 512, 1024) and t["grab0"] in (0, 256, 512, 1024)
    assert t["grab_shadow"] in (0, 64, 128)
    assert 0 < t["frame_ms"] <= 1.25 * t["default_ms"] and t["trials"] >= 52
    rd.render(gs, w.cam, pc, seed=2)
    assert same(rd.read(0), before)
    rd.set_schedule(1, 48, 1024)  # an explicit schedule: the same image
    rd.render(gs, w.cam, pc, seed=2)
    assert same(rd.read(0), before)
    rd.set_schedule(2, 16, 64, 0, 2048)  # camera rays with their own reservation
    rd.render(gs, w.cam, pc, seed=2)
    assert same(rd.read(0), before)
    rd.set_schedule(4, 8, 512, 3, 0, 16)  # shadow rays with their own (clamped to 16)
    rd.render(gs, w.cam, pc, seed=2)
    assert same(rd.read(0), before)


def test_batched_denoise_equals_per_environment(gpu):
    from kuafu_amd.denoiser import Denoiser
    w, d, cams = small_envs()
    gs = GpuScene(d)
    n_env, size = len(cams), w.env_height
    rd = Renderer(size, size * n_env)
    rd.render_envs(gs, cams, w.pc, seed=1)
    tg = [rd.read(k).copy() for k in range(3)]
    den = Denoiser(random_seed=7)
    den.apply(rd)  # one batch of n_env images
    out = rd.read(0)
    for e in range(n_env):
        r1 = Renderer(size, size)
        rows = slice(e * size, (e + 1) * size)
        for k in range(3):
            r1.write(k, tg[k][rows])
        den.apply(r1)
        assert same(out[rows], r1.read(0)), "environment %d" % e
    assert not same(out, tg[0])


def _tail_workloads():
    from kuafu_amd.api import camera_ubo, push_constants
    out = [scenes.config1(48, 40, spp=3), scenes.config2(width=48, height=27, spp=2, path_depth=8, n_inst=24, subdiv=2)]
    soft = scenes.tiny_scene(5, n_spheres=5, soft=True)  # inline shadow rays in the wavefront too
    cam = camera_ubo((0.0, 2.5, 5.0), (0, -0.4, -1.0), (0, 1, 0), 48, 40, fx=50.0, fy=50.0)
    out.append(scenes.Workload("soft", soft, cam, push_constants(spp=3, path_depth=6), 48, 40))
    alpha = scenes.tiny_scene(2, n_spheres=2)  # any-hit alpha test on path rays
    sv, si = scenes.uv_sphere(12, 8, 0.6)
    for a, pos in ((0.5, (-0.6, 0.7, 0.8)), (0.2, (0.7, 0.6, 1.2))):
        g = alpha.add_geometry(sv, si, scenes.material(diffuse=(0.3, 0.8, 0.4), alpha=a), opaque=False)
        alpha.add_instance(g, scenes.mat_from_rows(np.eye(3), pos))
    cam = camera_ubo((0.0, 2.0, 5.0), (0, -0.3, -1.0), (0, 1, 0), 48, 40, fx=50.0, fy=50.0)
    out.append(scenes.Workload("alpha", alpha, cam, push_constants(spp=4, path_depth=5), 48, 40))
    return out


def _assert_tail_equals_wavefront(w, tails, pairs=(1, 0)):
    gs = GpuScene(w.data)
    pc = w.pc
    pc.frame_count = 0
    ref = Renderer(w.width, w.height)
    st0 = ref.render(gs, w.cam, pc, seed=7, keep_samples=True)
    smp0 = ref.read_samples(pc.spp)
    tg0 = [ref.read(k) for k in range(3)]
    for tail in tails:
        for tp in pairs:
            name = "%s tail %d pairs %d" % (w.name, tail, tp)
            rd = Renderer(w.width, w.height)
            rd.set_schedule(0, 0, 0, tail, tail_pairs=tp)
            st = rd.render(gs, w.cam, pc, seed=7, keep_samples=True)
            assert same(rd.read_samples(pc.spp), smp0), name
            for k in range(3):
                assert same(rd.read(k), tg0[k]), "%s target %d" % (name, k)
            assert st.primary_rays == st0.primary_rays and st.shadow_rays == st0.shadow_rays, name
            # the wavefront also traces rays that Russian roulette then drops
            assert st.extension_rays == st0.extension_rays, name


def test_tail_kernel_equals_wavefront(gpu):
    """Bounces finished by a tail kernel give the wavefront's per-sample
    radiance, targets and ray counts bit for bit, whichever bounce it takes
    over from: k_tail2 (two lanes per path, the closest-hit ray of bounce b
    beside the shadow rays of bounce b-1; tail_pairs 1, scenes without
    random light perturbations) and k_tail (one thread per path, shadow rays
    inline; tail_pairs 0, and the soft-shadow scene either way)."""
    for w in _tail_workloads():
        _assert_tail_equals_wavefront(w, range(1, w.pc.path_depth + 1))


def test_tail_kernel_beyond_pair_capacity(gpu):
    """At bounce 1 of a 480x360 frame at 3 spp (518400 paths, more than half
    of them on the ground or a sphere) more paths are alive than k_tail2 has
    lane pairs for (a quarter of a launch's threads: 131072 on 256 CUs), so
    k_tail takes them although tail_pairs is 1."""
    _assert_tail_equals_wavefront(scenes.config1(480, 360, spp=3), (1, 2, 3), pairs=(1,))
