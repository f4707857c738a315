"""kf_trace_closest_noalloc (marked gpu): the allocation-free variant of
kf_trace_closest, called through the C ABI, returns the same hits bit for
bit on two calls in a row, the second with fewer rays (the scratch keeps the
first call's contents), and rejects batches above its scratch's maximum.
"""
import numpy as np
import pytest

from kuafu_amd import KuafuError, abi, scenes
from kuafu_amd.abi import check, lib, ptr
from kuafu_amd.api import GpuScene

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 18  # KF_TRACE_NOALLOC_MAX_RAYS


def _rays(rng, n, center, spread, box):
    r = np.zeros(n, abi.RAY_DTYPE)
    r["org"] = rng.uniform(-box, box, (n, 3))
    d = np.asarray(center) + rng.uniform(-spread, spread, (n, 3)) - r["org"]
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmin"] = 0.001
    r["tmax"] = 1e4
    return r


def _noalloc(gs, rays):
    rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
    hits = np.zeros(len(rays), abi.HIT_DTYPE)
    check(lib().kf_trace_closest_noalloc(gs.h, ptr(rays), len(rays), ptr(hits)))
    return hits


def test_noalloc_trace_equals_trace_closest(gpu):
    print("no-alloc trace: config1 scene, 2^18 then 3001 rays", flush=True)
    rng = np.random.default_rng(21)
    gs = GpuScene(scenes.config1().data)
    for n in (MAX_RAYS, 3001):
        rays = _rays(rng, n, (0, 0.5, 0), 3.0, 6.0)
        ref = gs.trace_closest(rays)
        got = _noalloc(gs, rays)
        assert (ref["inst"] != 0xFFFFFFFF).sum() > n // 10
        assert got.tobytes() == ref.tobytes(), n
    with pytest.raises(KuafuError, match="exceed"):
        _noalloc(gs, np.zeros(MAX_RAYS + 1, abi.RAY_DTYPE))
    assert _noalloc(gs, np.zeros(0, abi.RAY_DTYPE)).size == 0
