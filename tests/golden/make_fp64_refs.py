"""Write the FP64 oracle reference images used as converged targets.

* fp64_config1_1024spp.npy: the 256x256 spheres frame (config 1), 1024
  samples per pixel, FP64 oracle, seed 424242 (a stream no GPU test or bench
  frame uses), RGB as float16 (256, 256, 3).
* fp64_config2_rows_4096spp.npy: rows 15, 45, ..., 1065 (every 30th) of the
  1920x1080 icosphere frame (config 2), 4096 samples per pixel, FP64 oracle,
  seed 424242, RGB as float16 (36, 1920, 3).

float16 keeps these files small; its rounding (2^-11 relative) adds about
1e-7 to a relMSE.  Run from the repository root:
    python tests/golden/make_fp64_refs.py [config1|config2]
(about 1 and 8 minutes on 8 cores).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from kuafu_amd import scenes  # noqa: E402

SEED = 424242
C2_ROWS = list(range(15, 1080, 30))


def config1():
    w = scenes.config1(spp=1024)
    t = time.time()
    r = oracle.Scene(w.data).render(w.cam, w.pc, w.width, w.height, seed=SEED, precision=64, threads=os.cpu_count())
    np.save(os.path.join(HERE, "fp64_config1_1024spp.npy"), r["rgba"][..., :3].astype(np.float16))
    print("config1: %.0f s, %d rays" % (time.time() - t, int(r["counts"].sum())))


def config2():
    w = scenes.config2(spp=4096)
    o = oracle.Scene(w.data)
    out = np.zeros((len(C2_ROWS), w.width, 3), np.float32)
    t = time.time()
    for k, y in enumerate(C2_ROWS):
        r = o.render(w.cam, w.pc, w.width, w.height, seed=SEED, rows=(y, y + 1), precision=64,
                     threads=os.cpu_count())
        out[k] = r["rgba"][y, :, :3]
        print("row %d (%d/%d) %.0f s" % (y, k + 1, len(C2_ROWS), time.time() - t), flush=True)
    np.save(os.path.join(HERE, "fp64_config2_rows_4096spp.npy"), out.astype(np.float16))


if __name__ == "__main__":
    which = sys.argv[1:] or ["config1", "config2"]
    if "config1" in which:
        config1()
    if "config2" in which:
        config2()
