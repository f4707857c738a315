"""GPU parity at the benchmark scales and on the shading branches the small
frames leave out (marked gpu).

* Material textures (diffuse, metallic, roughness, transmission;
  PathTrace.rchit:339-360) and a cube-map environment (rmiss:14-27) on the
  config-1 scene, per sample against the FP32 oracle.
* The config-2 geometry at its full size (256 icosphere instances of
  subdivision 4, 1.31 M triangles) at a reduced resolution, per sample.
* The 1080p, 16 spp, depth-8 frame itself: per-sample parity on row bands at
  the top, middle and bottom of the frame, every primary ray counted,
  finite radiance, and the same image and ray counts whether its 33 M paths
  run as one wavefront batch or as four.
* The config-3 RGB camera (the IR cameras are in test_gpu_parity.py) and the
  IR cameras' FP32 single-channel output; the full active-light frame (720p,
  32 spp, depth 6, both IR cameras and the RGB camera) on row bands against
  the oracle, with the IR images' single-channel download.
* A converged GPU frame against the FP64 oracle's 1024-spp frame of the
  config-1 scene (tests/golden/fp64_config1_1024spp.npy): its relMSE must be
  the noise floor of the two estimates, measured from the per-pixel sample
  variance, with no excess a bias would add.
* The 128-environment frame at its real size (128 x 512x512, 8 spp, depth
  6, 11.9 M instanced triangles, 67 M paths in 4 batches): every camera ray
  counted, finite pixels, the rows of three environments equal bit for bit
  to renders of those environments alone and close to the oracle, and the
  batched denoise of all 128 images equal to denoising one of them alone.
"""
import os

import numpy as np
import pytest

import oracle
from kuafu_amd import abi, scenes
from kuafu_amd.api import GpuScene, Renderer
from test_gpu_parity import assert_frame_parity, render_both

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def textured_config1(width=64, height=64, spp=4):
    w = scenes.config1(width, height, spp=spp)
    s = w.data
    rng = np.random.default_rng(12)

    def tex(h, wd):
        t = rng.integers(0, 256, (h, wd, 4)).astype(np.uint8)
        t[..., 3] = 255
        return s.add_texture(t)

    # materials in config1's order: ground, red diffuse, specular, metal, glass, emissive
    s.materials[0]["diffuse_tex"] = tex(32, 32)
    s.materials[1]["diffuse_tex"] = tex(16, 8)
    rough = tex(8, 16)
    s.materials[2]["roughness_tex"] = rough
    s.materials[3]["metallic_tex"] = tex(8, 8)
    s.materials[3]["roughness_tex"] = rough
    s.materials[4]["transmission_tex"] = tex(4, 4)
    faces = []
    for _ in range(6):
        f = rng.integers(0, 256, (16, 16, 4)).astype(np.uint8)
        f[..., 3] = 255
        faces.append(f)
    s.env_kind = abi.KF_ENV_CUBEMAP
    s.env_faces = faces
    w.pc.use_env_map = 1
    return w


def test_render_parity_textured_materials_and_cubemap(gpu):
    print("textured config1 64x64 4spp + cube map", flush=True)
    w = textured_config1()
    (res,) = render_both(w)
    assert_frame_parity(*res)
    # the textures and the sky reach the image: albedo varies on the ground
    alb = res[3][..., :3]
    assert len(np.unique(alb[-8:, :, 0])) > 8


def test_cubemap_scene_teardown_leaves_no_error(gpu):
    # the cube-map faces' records live inside the texture table: destroying
    # the scene must free each allocation once and leave no error behind for
    # the next call to find
    w = textured_config1(16, 16, spp=1)
    for _ in range(2):
        gs = GpuScene(w.data)
        Renderer(16, 16).render(gs, w.cam, w.pc)
        gs.close()
    gs = GpuScene(scenes.config1(16, 16, spp=1).data)
    gs.close()


def test_render_parity_config2_full_geometry(gpu):
    print("config2 geometry (256 x 5120 triangles) at 96x54, 2 spp, depth 8", flush=True)
    w = scenes.config2(width=96, height=54, spp=2)
    assert w.data.n_triangles() == 256 * 5120 + 2
    (res,) = render_both(w)
    assert_frame_parity(*res)


def test_config2_1080p_frame(gpu):
    print("config2 1920x1080 16 spp depth 8: bands against the oracle", flush=True)
    w = scenes.config2()
    gs = GpuScene(w.data)
    rd = Renderer(w.width, w.height)
    st = rd.render(gs, w.cam, w.pc, seed=5, samples_per_batch=16, keep_samples=True)
    smp = rd.read_samples(w.pc.spp)
    rgba = rd.read(0)
    n_px = w.width * w.height
    assert st.primary_rays == n_px * w.pc.spp
    # about half the camera rays leave the scene for the sky
    assert st.extension_rays > 0.3 * st.primary_rays and st.shadow_rays > 0
    assert (rgba[..., 3] == 1).all()
    fin = np.isfinite(rgba[..., :3]).all(-1).mean()
    print("finite pixels: %.6f, rays: %d primary, %d extension, %d shadow"
          % (fin, st.primary_rays, st.extension_rays, st.shadow_rays))
    assert fin > 0.9999
    o = oracle.Scene(w.data)
    for y0, y1 in ((0, 2), (539, 542), (1077, 1080)):
        ref = o.render(w.cam, w.pc, w.width, w.height, seed=5, rows=(y0, y1), threads=16, keep_samples=True)
        rs, gs_ = ref["samples"][:, y0:y1], smp[:, y0:y1]
        close = (np.abs(gs_ - rs) <= 1e-4 * np.maximum(1.0, np.abs(rs))) | (np.isnan(gs_) & np.isnan(rs))
        print("rows %d-%d: per-sample within 1e-4 %.6f" % (y0, y1, close.all(-1).mean()))
        assert close.all(-1).mean() >= 0.999
        ok = np.abs(rgba[y0:y1] - ref["rgba"][y0:y1]) <= 1e-4 * np.maximum(1.0, np.abs(ref["rgba"][y0:y1]))
        assert ok.mean() > 0.999
    del smp
    # the same frame as four wavefront batches of 4 samples: same bits, same rays
    st4 = rd.render(gs, w.cam, w.pc, seed=5, samples_per_batch=4)
    assert (rd.read(0) == rgba).all() or np.array_equal(rd.read(0), rgba, equal_nan=True)
    assert (st4.primary_rays, st4.extension_rays, st4.shadow_rays) == \
        (st.primary_rays, st.extension_rays, st.shadow_rays)


def test_render_parity_config3_rgb_camera(gpu):
    print("config3 rgb camera 96x54", flush=True)
    w = scenes.config3(width=96, height=54, spp=2, path_depth=4, camera="rgb")
    (res,) = render_both(w)
    assert_frame_parity(*res)


def test_ir_camera_single_channel_output(gpu):
    # the IR cameras' FP32 single-channel image is (r + g + b) / 3 of the
    # rgba target, in FP32
    w = scenes.config3(width=96, height=54, spp=2, path_depth=4, camera="left")
    rd = Renderer(w.width, w.height)
    rd.render(GpuScene(w.data), w.cam, w.pc, seed=2)
    rgba = rd.read(0)
    gray = rd.download_gray32f()
    exp = (rgba[..., 0] + rgba[..., 1] + rgba[..., 2]) / np.float32(3.0)
    assert gray.dtype == np.float32 and gray.shape == (54, 96)
    assert np.array_equal(gray, exp, equal_nan=True)
    assert gray.max() > 0


def test_converged_frame_against_fp64_oracle(gpu):
    print("config1 4096 spp against the FP64 1024-spp fixture", flush=True)
    ref = np.load(os.path.join(GOLDEN, "fp64_config1_1024spp.npy")).astype(np.float32)
    w = scenes.config1()
    gs = GpuScene(w.data)
    rd = Renderer(w.width, w.height)
    pc = w.pc
    pc.spp = 4096
    rd.render(gs, w.cam, pc, seed=0, samples_per_batch=64)
    img = rd.read(0)[..., :3].copy()
    e = oracle.relmse(img, ref)
    # noise floor: per-pixel variance of one sample (from a 64-spp frame)
    # divided by each estimate's sample count
    pc.spp = 64
    rd.render(gs, w.cam, pc, seed=77, keep_samples=True)
    smp = rd.read_samples(64).astype(np.float64)
    var = smp.var(0, ddof=1)
    floor = float(np.mean(var * (1 / 1024 + 1 / 4096) / (ref.astype(np.float64) ** 2 + 1e-2)))
    print("relMSE(GPU 4096 spp, FP64 1024 spp) = %.3e; noise floor %.3e" % (e, floor))
    assert e < 1.5 * floor + 2e-4, (e, floor)
    assert e < 2e-3


def test_config5_full_scale(gpu):
    from kuafu_amd.denoiser import Denoiser, WEIGHTS
    print("config5 128 environments of 512x512, 8 spp, depth 6", flush=True)
    w = scenes.config5()
    d, pc, size, n_env = w.data, w.pc, w.env_height, w.n_env
    pc.frame_count = 0
    gs = GpuScene(d)
    rd = Renderer(size, size * n_env)
    st = rd.render_envs(gs, w.cam, pc, seed=11)
    assert st.primary_rays == n_env * size * size * pc.spp
    assert st.extension_rays > 0 and st.shadow_rays > 0
    tg = [rd.read(k) for k in range(3)]
    fin = np.isfinite(tg[0][..., :3]).all(-1).mean()
    print("finite pixels %.6f; rays %d primary, %d extension, %d shadow"
          % (fin, st.primary_rays, st.extension_rays, st.shadow_rays), flush=True)
    assert fin > 0.9999
    for e in (0, 77, n_env - 1):
        sub = d.environment(e)
        r1 = Renderer(size, size)
        r1.render(GpuScene(sub), w.cam[e], pc, seed=11)
        rows = slice(e * size, (e + 1) * size)
        for k in range(3):
            got, alone = tg[k][rows], r1.read(k)
            assert ((got == alone) | (np.isnan(got) & np.isnan(alone))).all(), "environment %d target %d" % (e, k)
        # two rows against the oracle on the environment's own scene
        o = oracle.Scene(sub)
        ref = o.render(w.cam[e], pc, size, size, seed=11, rows=(255, 257), threads=16)
        got = tg[0][rows][255:257]
        ok = np.abs(got - ref["rgba"][255:257]) <= 1e-4 * np.maximum(1.0, np.abs(ref["rgba"][255:257]))
        print("environment %d rows 255-256 within 1e-4 of the oracle: %.5f" % (e, ok.mean()), flush=True)
        assert ok.mean() > 0.999
    # batched denoise of the 128 stacked images: environment 77 as if alone
    den = Denoiser(weights=WEIGHTS)
    den.apply(rd)
    out = rd.read(0)
    r1 = Renderer(size, size)
    rows = slice(77 * size, 78 * size)
    for k in range(3):
        r1.write(k, tg[k][rows])
    den.apply(r1)
    assert (out[rows] == r1.read(0)).all()
    assert np.isfinite(out[..., :3]).all()


def test_config3_full_frame(gpu):
    print("config3 1280x720, 32 spp, depth 6: IR left, IR right, RGB", flush=True)
    for cam in ("left", "right", "rgb"):
        w = scenes.config3(camera=cam)
        gs = GpuScene(w.data)
        rd = Renderer(w.width, w.height)
        st = rd.render(gs, w.cam, w.pc, seed=4)
        assert st.primary_rays == w.width * w.height * w.pc.spp
        rgba = rd.read(0)
        assert np.isfinite(rgba[..., :3]).all()
        o = oracle.Scene(w.data)
        for y0, y1 in ((0, 1), (359, 361), (718, 720)):
            ref = o.render(w.cam, w.pc, w.width, w.height, seed=4, rows=(y0, y1), threads=16)["rgba"][y0:y1]
            ok = np.abs(rgba[y0:y1] - ref) <= 1e-4 * np.maximum(1.0, np.abs(ref))
            print("%s rows %d-%d within 1e-4 of the oracle: %.5f" % (cam, y0, y1, ok.mean()), flush=True)
            assert ok.mean() > 0.999
        if cam != "rgb":
            gray = rd.download_gray32f()
            assert gray.shape == (w.height, w.width)
            assert (gray == (rgba[..., 0] + rgba[..., 1] + rgba[..., 2]) / np.float32(3.0)).all()
