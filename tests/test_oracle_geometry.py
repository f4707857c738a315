"""Pins of the oracle's ray/triangle test and traversal (CPU only).

Watertight test (Woo, Benthin, Wald 2013): checked against exact float64
intersection, on shared edges and vertices (a ray through an edge or vertex
of a closed fan must hit), on degenerate triangles, and for the facing rule
front <=> dot(d, cross(v1 - v0, v2 - v0)) < 0.  The oracle's BVH traversal is
checked against brute force over every triangle of every instance.
"""
import numpy as np
import pytest

import oracle
from kuafu_amd import abi, scenes
from kuafu_amd.api import SceneData, mat_from_rows, material


def exact_hit(o, d, v0, v1, v2):
    o, d, v0, v1, v2 = (np.asarray(a, np.float64) for a in (o, d, v0, v1, v2))
    e1, e2 = v1 - v0, v2 - v0
    M = np.stack([-d, e1, e2], 1)
    if abs(np.linalg.det(M)) < 1e-300:
        return None
    t, b1, b2 = np.linalg.solve(M, o - v0)
    return t, b1, b2


def test_random_hits_match_exact():
    rng = np.random.default_rng(0)
    n_hit = 0
    for _ in range(3000):
        v = rng.uniform(-1, 1, (3, 3))
        o = rng.uniform(-3, 3, 3)
        target = v[0] * 0.2 + v[1] * 0.3 + v[2] * 0.5 + rng.normal(0, 0.3, 3)
        d = target - o
        hit, t, b1, b2, back = oracle.intersect_triangle(o, d, 0.0, 1e30, *v)
        ex = exact_hit(o.astype(np.float32), d.astype(np.float32), *v.astype(np.float32))
        if ex is None:
            continue
        te, b1e, b2e = ex
        inside = b1e > 1e-4 and b2e > 1e-4 and b1e + b2e < 1 - 1e-4 and te > 1e-4
        outside = b1e < -1e-4 or b2e < -1e-4 or b1e + b2e > 1 + 1e-4 or te < -1e-4
        if inside:
            assert hit
            n_hit += 1
            assert abs(t - te) <= 1e-5 * abs(te) + 1e-6
            assert abs(b1 - b1e) < 1e-4 and abs(b2 - b2e) < 1e-4
            n = np.cross(v[1] - v[0], v[2] - v[0])
            assert back == (1 if np.dot(d, n) > 0 else 0)
        elif outside:
            assert not hit
    assert n_hit > 500


def test_shared_edge_and_vertex_are_watertight():
    rng = np.random.default_rng(2)
    # fan of 8 triangles around a centre vertex in the z=0 plane
    c = np.array([0.1, -0.2, 0.0])
    ring = [np.array([np.cos(a), np.sin(a), 0.0]) for a in np.linspace(0, 2 * np.pi, 9)[:-1]]
    tris = [(c, ring[i], ring[(i + 1) % 8]) for i in range(8)]
    for _ in range(200):
        o = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(1, 3)])
        # aim exactly at the shared centre vertex, then at points on shared edges
        for p in [c] + [c + s * (ring[i] - c) for i in range(8) for s in (0.25, 0.5)]:
            d = (p - o).astype(np.float32)
            hits = [oracle.intersect_triangle(o, d, 0.0, 1e30, *t)[0] for t in tris]
            assert any(hits), "ray through a shared edge/vertex fell through"


def test_degenerate_triangle_never_hits():
    v0 = np.array([0, 0, 0.0]); v1 = np.array([1, 1, 1.0]); v2 = np.array([2, 2, 2.0])
    for d in ([1, 1, 1], [0.5, 0.7, 1.0], [-1, 1, 0.3]):
        assert not oracle.intersect_triangle([-1, -1, -1], d, 0.0, 1e30, v0, v1, v2)[0]


def test_tmin_tmax_inclusive_bounds():
    v0, v1, v2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    o, d = [0.25, 0.25, 2.0], [0, 0, -1.0]
    hit, t, *_ = oracle.intersect_triangle(o, d, 0.0, 1e30, v0, v1, v2)
    assert hit and t == 2.0
    assert oracle.intersect_triangle(o, d, 2.0, 2.0, v0, v1, v2)[0]
    assert not oracle.intersect_triangle(o, d, 2.0000002, 3.0, v0, v1, v2)[0]
    assert not oracle.intersect_triangle(o, d, 0.0, 1.9999999, v0, v1, v2)[0]


def random_rays(rng, n, lo, hi, tmax=1e4):
    r = np.zeros(n, abi.RAY_DTYPE)
    r["org"] = rng.uniform(lo, hi, (n, 3))
    tgt = rng.uniform(lo, hi, (n, 3)) * 0.5
    d = tgt - r["org"]
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmin"] = 0.001
    r["tmax"] = tmax
    return r


def hits_equal(a, b):
    assert (a["inst"] == b["inst"]).all()
    m = a["inst"] != 0xFFFFFFFF
    for k in ("prim", "back"):
        assert (a[k][m] == b[k][m]).all()
    for k in ("t", "b1", "b2"):
        assert (a[k][m] == b[k][m]).all()


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_bvh_equals_brute_force(seed):
    s = scenes.tiny_scene(seed, n_spheres=4)
    o = oracle.Scene(s)
    rng = np.random.default_rng(seed)
    rays = random_rays(rng, 4000, -4, 4)
    hits_equal(o.trace_closest(rays), o.trace_closest(rays, brute=True))
    rays["tmax"] = rng.uniform(0.5, 6, len(rays)).astype(np.float32)
    assert (o.trace_any(rays) == o.trace_any(rays, brute=True)).all()


def test_world_to_object_inverts_and_matches_product():
    from kuafu_amd import lib
    rng = np.random.default_rng(3)
    for _ in range(50):
        R = scenes.random_rotation(rng) * rng.uniform(0.1, 3.0, 3)
        T = mat_from_rows(R, rng.uniform(-10, 10, 3))
        w = oracle.world_to_object(T)
        prod = np.zeros(12, np.float32)
        assert lib().kf_world_to_object(T.ravel().ctypes.data, prod.ctypes.data) == 0
        assert (w.view(np.uint32) == prod.view(np.uint32)).all()
        A = np.eye(4)
        A[:3, :] = np.asarray(T, np.float64).T[:3, :]
        W = np.eye(4)
        W[:3, :] = w.reshape(3, 4)
        assert np.allclose(W @ A, np.eye(4), atol=1e-5)
