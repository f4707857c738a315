"""GPU checks of the denoiser's MFMA kernels against the FP64 reference (marked gpu).

* One 3x3 convolution layer: forward (bias + ReLU, BF16 output), data
  gradient (ReLU-masked, BF16), weight and bias gradients (FP32) against
  numpy FP64 on the same BF16 operands.
* The whole U-Net: loss and the gradient of all 1.8 M parameters against
  oracle/unet_ref.py (FP64, BF16 storage emulated); the reference itself is
  pinned by finite differences in tests/test_unet_ref.py.
* Training reduces the loss; weights survive save/load.
"""
import numpy as np
import pytest

from kuafu_amd.denoiser import Denoiser, conv_backward, conv_forward, from_bf16_bits, to_bf16_bits
from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def bf(x):
    return from_bf16_bits(to_bf16_bits(x)).astype(np.float64)


# The kernel works on 32 x 8-pixel tiles; a tile whose 34 x 10 halo lies
# inside the image takes a path without bounds tests, so the larger shapes
# (24 x 72, 20 x 100) include such interior tiles besides edge tiles.
@pytest.mark.parametrize("cin,cout,h,w", [(32, 32, 8, 8), (32, 64, 6, 10), (96, 32, 5, 7), (288, 128, 4, 4),
                                          (64, 96, 9, 3), (32, 3, 7, 5), (160, 160, 3, 3), (32, 32, 24, 72),
                                          (64, 32, 20, 100)])
def test_conv_forward(gpu, cin, cout, h, w):
    rng = np.random.default_rng(cin + cout)
    x = bf(rng.normal(0, 1, (2, h, w, cin)))
    wt = bf(rng.normal(0, 0.1, (9, cout, cin)))
    b = rng.normal(0, 0.1, cout).astype(np.float32)
    y = conv_forward(x, wt, b, relu=True)
    ref = U.round_bf16(np.maximum(U.conv_fwd(x, wt, b.astype(np.float64)), 0))
    diff = np.abs(y - ref)
    assert (diff == 0).mean() > 0.97
    assert (diff <= 0.01 * np.abs(ref) + 1e-6).all()


@pytest.mark.parametrize("cin,cout,h,w", [(32, 32, 8, 8), (64, 32, 6, 10), (96, 64, 5, 7), (288, 128, 4, 4),
                                          (32, 160, 3, 5), (64, 32, 24, 72)])
def test_conv_backward(gpu, cin, cout, h, w):
    rng = np.random.default_rng(3 * cin + cout)
    x = bf(rng.normal(0, 1, (2, h, w, cin)))
    wt = bf(rng.normal(0, 0.1, (9, cout, cin)))
    dy = bf(rng.normal(0, 1, (2, h, w, cout)))
    dx, dw, db = conv_backward(x, wt, dy, mask_by_x=True)
    rdx, rdw, rdb = U.conv_bwd(x, wt, dy)
    rdx = U.round_bf16(rdx * (x > 0))
    diff = np.abs(dx - rdx)
    assert (diff == 0).mean() > 0.97
    assert (diff <= 0.01 * np.abs(rdx) + 1e-6).all()
    assert np.linalg.norm(dw - rdw) <= 1e-5 * np.linalg.norm(rdw)
    assert np.allclose(db, rdb, rtol=1e-5, atol=1e-4)


def he_params(seed):
    rng = np.random.default_rng(seed)
    sl, n = U.param_slices()
    p = np.zeros(n, np.float32)
    for k, (w, b, ci, co) in enumerate(sl):
        W = rng.normal(0, np.sqrt(2.0 / (9 * ci)), (9, co, ci))
        if k == len(sl) - 1:
            W[:, 3:, :] = 0.0
        p[w[0]:w[1]] = W.ravel()
        p[b[0]:b[1]] = rng.normal(0, 0.02, co) if k < len(sl) - 1 else np.r_[rng.normal(0, 0.02, 3), np.zeros(co - 3)]
    return p


def batch(seed, n=2, h=32, w=32):
    rng = np.random.default_rng(seed)
    inp = np.zeros((n, h, w, 9), np.float32)
    inp[..., 0:3] = np.abs(rng.normal(0.3, 0.6, (n, h, w, 3)))
    inp[..., 3:6] = rng.uniform(0, 1 / 3.141592, (n, h, w, 3))
    inp[..., 6:9] = rng.normal(0, 1, (n, h, w, 3))
    tgt = np.abs(rng.normal(0.3, 0.3, (n, h, w, 3))).astype(np.float32)
    return inp, tgt


def test_network_gradients_match_reference(gpu):
    d = Denoiser(random_seed=0)
    p = he_params(1)
    d.set_params(p)
    # One 128 x 128 image: 256 pixels at level 3 and 64 at level 4.  Targets
    # above every output make the loss gradient one-signed, so the weight
    # gradients of the deep layers add up over pixels instead of cancelling,
    # and the BF16 rounding flips between the GPU's FP32 accumulation and the
    # FP64 reference stay small against them.
    inp, tgt = batch(2, n=1, h=128, w=128)
    tgt = (tgt + 2.0).astype(np.float32)
    loss, g = d.gradients(inp, tgt)
    net = U.Net(p, bf16=True)
    s = net.forward(inp)
    rl, dl = net.loss(s, tgt)
    rg = net.backward(s, dl)
    assert abs(loss - rl) <= 1e-3 * abs(rl), (loss, rl)
    sl, _ = U.param_slices()
    # Tolerance: activations are stored in BF16; a value within one FP32
    # rounding of a BF16 rounding boundary rounds one way on the GPU (FP32
    # MFMA accumulation) and the other in the FP64 reference.  On this batch
    # such flips leave at most ~1 % relative error in any tensor (median
    # 0.4 %); a wrong tap, channel or mask gives errors of order 1.
    errs = []
    for k, (w, b, ci, co) in enumerate(sl):
        for a, z in ((w[0], w[1]), (b[0], b[1])):
            ref = rg[a:z]
            if np.linalg.norm(ref) == 0:
                assert np.abs(g[a:z]).max() == 0
                continue
            err = np.linalg.norm(g[a:z] - ref) / np.linalg.norm(ref)
            errs.append(err)
            assert err < 2e-2, (k, a, err)
    print("per-tensor relative gradient error: worst %.2e median %.2e" % (max(errs), np.median(errs)))
    assert np.median(errs) < 1e-2
    # forward output against the reference
    out = d.forward(inp)
    ro = net.output(s)
    assert np.abs(out - ro).max() <= 1e-2 * np.abs(ro).max()


def test_network_forward_interior_tiles(gpu):
    # 96 x 160: interior tiles at levels 0 and 1, where the fused pooling
    # (encoder) and upsampling (decoder) run
    d = Denoiser(random_seed=0)
    p = he_params(5)
    d.set_params(p)
    inp, _ = batch(6, n=1, h=96, w=160)
    net = U.Net(p, bf16=True)
    ro = net.output(net.forward(inp))
    out = d.forward(inp)
    assert np.abs(out - ro).max() <= 1e-2 * np.abs(ro).max()


def test_training_reduces_loss_and_save_load(gpu, tmp_path):
    d = Denoiser(random_seed=3)
    inp, tgt = batch(5, n=4, h=32, w=48)
    losses = [d.train_step(inp, tgt, 2e-3) for _ in range(40)]
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.7 * losses[0], (losses[0], losses[-1])
    f = tmp_path / "w.kfw"
    d.save(f)
    e = Denoiser(weights=f)
    assert np.array_equal(e.params(), bf(d.params()).astype(np.float32)) or \
        np.allclose(e.params(), d.params(), rtol=1e-2, atol=1e-6)
    assert np.allclose(e.forward(inp), Denoiser(weights=f).forward(inp))


@pytest.mark.parametrize("width,height", [(64, 48), (160, 96)])
def test_denoise_stage_on_render_targets(gpu, width, height):
    # the render-target path (first convolution reading the float4 targets)
    # against the record path; 160 x 96 has interior tiles at levels 0 and 1
    from kuafu_amd import GpuScene, Renderer, scenes
    w = scenes.config1(width, height, spp=1)
    r = Renderer(w.width, w.height)
    r.render(GpuScene(w.data), w.cam, w.pc)
    before = r.read(0)
    d = Denoiser(random_seed=4)
    d.apply(r)
    after = r.read(0)
    assert np.isfinite(after).all() and (after[..., 3] == 1).all()
    # against the record path on the same inputs (16-aligned crop of the frame)
    hc, wc = height // 16 * 16, width // 16 * 16
    rec = np.concatenate([before[..., :3], r.read(1)[..., :3], r.read(2)[..., :3]], -1)[None, :hc, :wc]
    ref = d.forward(np.ascontiguousarray(rec))
    assert np.abs(ref[0] - after[:hc, :wc, :3]).max() <= 1e-3 * max(1.0, np.abs(ref).max())


def test_denoise_stage_ragged_raster(gpu):
    # a raster whose sides are not multiples of 16: the render-target path
    # pads it to 80 x 112 by repeating edge pixels (zeros beyond); the record
    # path on the same padding done here must give the same image
    from kuafu_amd import GpuScene, Renderer, scenes
    width, height = 100, 70
    w = scenes.config1(width, height, spp=1)
    r = Renderer(w.width, w.height)
    r.render(GpuScene(w.data), w.cam, w.pc)
    before = r.read(0)
    d = Denoiser(random_seed=5)
    d.apply(r)
    after = r.read(0)
    assert np.isfinite(after).all() and (after[..., 3] == 1).all()
    rec = np.concatenate([before[..., :3], r.read(1)[..., :3], r.read(2)[..., :3]], -1)[None]
    hp, wp = -(-height // 16) * 16, -(-width // 16) * 16
    rec = np.pad(rec, ((0, 0), (0, hp - height), (0, wp - width), (0, 0)), mode="edge")
    ref = d.forward(np.ascontiguousarray(rec, np.float32))[0, :height, :width]
    assert np.abs(ref - after[..., :3]).max() <= 1e-3 * max(1.0, np.abs(ref).max())


def test_trained_network_gain(gpu):
    # The shipped weights (kuafu_amd/weights/unet.kfw, trained on seeds
    # 0-1999 of the training generator) on held-out scenes of that generator
    # and on the config-1 frame against the FP64 oracle's 1024-spp image: the
    # denoised frame's relMSE must beat the noisy input's by a wide margin.
    import os

    import oracle
    from kuafu_amd import GpuScene, Renderer, scenes
    from kuafu_amd.denoiser import WEIGHTS
    assert os.path.exists(WEIGHTS)
    d = Denoiser(weights=WEIGHTS)
    cases = []
    for seed in (3001, 3002, 3003):
        w = scenes.training_scene(seed)
        gs = GpuScene(w.data)
        rd = Renderer(w.width, w.height)
        pc = w.pc
        pc.spp = 2048
        rd.render(gs, w.cam, pc, seed=999, samples_per_batch=64)
        cases.append(("heldout%d" % seed, w, gs, rd, rd.read(0)[..., :3].copy()))
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fp64_config1_1024spp.npy")
    w = scenes.config1()
    gs = GpuScene(w.data)
    cases.append(("config1_fp64", w, gs, Renderer(w.width, w.height), np.load(golden).astype(np.float32)))
    for name, w, gs, rd, ref in cases:
        for spp, factor in ((1, 10.0), (4, 5.0)):
            pc = w.pc
            pc.spp = spp
            pc.frame_count = 0
            rd.render(gs, w.cam, pc, seed=1)
            noisy = rd.read(0)[..., :3].copy()
            d.apply(rd)
            den = rd.read(0)[..., :3]
            en, ed = oracle.relmse(noisy, ref), oracle.relmse(den, ref)
            print("%s %d spp: relMSE noisy %.4f -> denoised %.4f (%.1fx)" % (name, spp, en, ed, en / ed))
            assert np.isfinite(den).all()
            assert ed * factor < en, (name, spp, en, ed)
