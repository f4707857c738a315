"""Train the denoiser U-Net on images this tracer renders itself.

For each seeded training scene (kuafu_amd.scenes.training_scene) the GPU path
tracer renders a high-spp target and low-spp inputs (noisy RGB + albedo +
normal AOVs at 1, 2, 4 and 8 spp).  Training runs the hand-written BF16 MFMA
forward/backward kernels with Adam (kf_denoiser_train_step) on random crops
with flips; held-out scenes report relMSE of the noisy input and of the
denoised output against the target.

  python -m kuafu_amd.train_denoiser --out weights.kfw

The defaults are the reference configuration's training set: 2000 seeded
scenes of 256x256, 1024-spp targets, 256x256 crops (here whole frames, with
flips).  Inputs are kept in FP16 on the host (2008 scenes x 4 levels x 9
channels = 9.5 GB), converted to FP32 per batch.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from kuafu_amd import GpuScene, Renderer, lib, scenes  # noqa: E402
from kuafu_amd.abi import check  # noqa: E402
from kuafu_amd.denoiser import Denoiser  # noqa: E402

SPP_LEVELS = (1, 2, 4, 8)


def relmse(x, y):
    return float(np.mean((x - y) ** 2 / (y ** 2 + 1e-2)))


def render_pair(rd, seed, size, target_spp):
    w = scenes.training_scene(seed, size, size)
    gs = GpuScene(w.data)
    pc = w.pc
    pc.frame_count = 0
    pc.spp = target_spp
    rd.render(gs, w.cam, pc, seed=7919 * seed + 1)
    target = rd.read(0)[..., :3].copy()
    inputs = []
    for spp in SPP_LEVELS:
        pc.spp = spp
        rd.render(gs, w.cam, pc, seed=104729 * seed + spp)
        inputs.append(np.concatenate([rd.read(0)[..., :3], rd.read(1)[..., :3], rd.read(2)[..., :3]], -1))
    gs.close()
    return np.stack(inputs).astype(np.float16), target.astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--heldout", type=int, default=8)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--schedule", default="256,16,256,5",
                    help="wavefront schedule of the renders: samples per batch, refill, reservation, tail bounce")
    ap.add_argument("--iters", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init", default=None, help="start from a weight file")
    ap.add_argument("--out", required=True)
    ap.add_argument("--log", default=None)
    args = ap.parse_args()
    check(lib().kf_init(0))
    t0 = time.time()
    n = args.scenes + args.heldout
    X = np.zeros((n, len(SPP_LEVELS), args.size, args.size, 9), np.float16)
    Y = np.zeros((n, args.size, args.size, 3), np.float32)
    rd = Renderer(args.size, args.size)
    rd.set_schedule(*[int(v) for v in args.schedule.split(",")])
    for i in range(n):
        X[i], Y[i] = render_pair(rd, i, args.size, args.target_spp)
        if i % 50 == 0:
            print("dataset %d/%d %.1fs" % (i, n, time.time() - t0), flush=True)
    print("dataset done %.1fs" % (time.time() - t0), flush=True)
    d = Denoiser(weights=args.init) if args.init else Denoiser(random_seed=1)
    rng = np.random.default_rng(0)
    c = args.crop
    log = []
    t1 = time.time()
    for it in range(args.iters):
        lr = args.lr * 0.5 * (1 + math.cos(math.pi * it / args.iters))
        si = rng.integers(0, args.scenes, args.batch)
        li = rng.integers(0, len(SPP_LEVELS), args.batch)
        ys = rng.integers(0, args.size - c + 1, args.batch)
        xs = rng.integers(0, args.size - c + 1, args.batch)
        bi = np.stack([X[s, l, y:y + c, x:x + c] for s, l, y, x in zip(si, li, ys, xs)]).astype(np.float32)
        bt = np.stack([Y[s, y:y + c, x:x + c] for s, y, x in zip(si, ys, xs)])
        if rng.uniform() < 0.5:
            bi, bt = bi[:, :, ::-1], bt[:, :, ::-1]
        if rng.uniform() < 0.5:
            bi, bt = bi[:, ::-1], bt[:, ::-1]
        loss = d.train_step(np.ascontiguousarray(bi), np.ascontiguousarray(bt), lr)
        if it % 100 == 0 or it == args.iters - 1:
            print("iter %d loss %.4f lr %.2e %.1fs" % (it, loss, lr, time.time() - t1), flush=True)
            log.append((it, loss))
    train_s = time.time() - t1
    d.save(args.out)
    # held-out evaluation
    ev = {}
    for li, spp in enumerate(SPP_LEVELS):
        rn, rd_ = [], []
        for s in range(args.scenes, n):
            xs_ = X[s, li].astype(np.float32)
            out = d.forward(xs_[None])[0]
            rn.append(relmse(xs_[..., :3], Y[s]))
            rd_.append(relmse(out, Y[s]))
        ev[str(spp)] = {"noisy_relmse": float(np.mean(rn)), "denoised_relmse": float(np.mean(rd_))}
    res = {"scenes": args.scenes, "heldout": args.heldout, "size": args.size, "target_spp": args.target_spp,
           "iters": args.iters, "batch": args.batch, "crop": c, "train_seconds": round(train_s, 1),
           "ms_per_iter": round(1e3 * train_s / max(args.iters, 1), 2), "final_loss": log[-1][1], "heldout_eval": ev}
    print(json.dumps(res))
    if args.log:
        with open(args.log, "w") as f:
            json.dump({"result": res, "loss_curve": log}, f)


if __name__ == "__main__":
    main()
