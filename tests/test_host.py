"""Host-side checks (CPU only): camera matrices, sRGB readback bytes, scene
generators, the C ABI surface and its input validation."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import oracle
from kuafu_amd import KuafuError, abi, lib, scenes
from kuafu_amd.api import GpuScene, SceneData, camera_ubo, look_at, material

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def m4(a):
    """float[16] column-major -> math-layout 4x4."""
    return np.asarray(list(a), np.float64).reshape(4, 4).T


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "kuafu_amd.h")).read()
    names = set(re.findall(r"\b(kf_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 20
    L = C.CDLL(abi.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert {n for n, _, _ in abi.EXPORTS} == names


def test_camera_ubo_pinhole():
    eye = np.array([1.0, 2.0, 3.0])
    front = np.array([0.3, -0.2, -1.0])
    front /= np.linalg.norm(front)
    W, H, fx, fy, cx, cy = 640, 480, 500.0, 520.0, 300.0, 250.0
    u = camera_ubo(eye, front, (0, 1, 0), W, H, fx=fx, fy=fy, cx=cx, cy=cy)
    V, P, Vi, Pi = m4(u.view), m4(u.proj), m4(u.view_inverse), m4(u.proj_inverse)
    assert np.allclose(V @ Vi, np.eye(4), atol=1e-5) and np.allclose(P @ Pi, np.eye(4), atol=1e-4)
    assert np.allclose(V @ np.r_[eye, 1], [0, 0, 0, 1], atol=1e-5)
    assert np.allclose(V[:3, :3] @ front, [0, 0, -1], atol=1e-6)
    # a world point in view projects to pixel (cx + fx x/-z, cy + fy (-y)/-z) (camera.cpp:111-128)
    rng = np.random.default_rng(0)
    for _ in range(20):
        pv = np.r_[rng.uniform(-1, 1, 2), -rng.uniform(1, 5), 1.0]
        clip = P @ pv
        ndc = clip[:3] / clip[3]
        px = (ndc[0] + 1) / 2 * W
        py = (ndc[1] + 1) / 2 * H
        assert abs(px - (cx + fx * pv[0] / -pv[2])) < 1e-3
        assert abs(py - (cy - fy * pv[1] / -pv[2])) < 1e-3
    # the raygen direction through the principal point is the front vector (rgen:45-56)
    d = np.array([2 * cx / W - 1, 2 * cy / H - 1, 1, 1])
    t = Pi @ d
    dirw = Vi @ np.r_[t[:3] / np.linalg.norm(t[:3]), 0]
    assert np.allclose(dirw[:3], front, atol=1e-5)
    assert u.position[3] == 0.0 and u.front[3] == 5.0


def test_active_light_fill_matches_scene_upload():
    s = SceneData()
    view = look_at((1, 2, 3), (0, 0, 0), (0, 0, 1))
    assert s.add_active_light(view, (0.5, 0.6, 0.7), fov=math.radians(40), strength=9.0, softness=0.0, tex_id=3)
    al = s.active_lights
    eye = np.array([1, 2, 3.0])
    assert np.allclose(list(al.position[0])[:3], eye, atol=1e-5)
    fr = -eye / np.linalg.norm(eye)
    assert np.allclose(list(al.front[0])[:3], fr, atol=1e-5) and al.front[0][3] == 1.0
    P = m4(al.proj[0])
    assert abs(P[1, 1] - 1 / math.tan(math.radians(20))) < 1e-4 and P[3, 2] == -1.0
    assert list(al.sftp[0]) == [0.0, np.float32(math.radians(40)), 3.0, 0.0]


def test_to_bgra8_srgb_bytes():
    x = np.concatenate([np.linspace(-0.1, 1.2, 20001), [np.nan, 0.0, 1.0, 0.0031308, 0.5]]).astype(np.float32)
    n = len(x) - len(x) % 4
    rgba = x[:n].reshape(-1, 4)
    out = oracle.to_bgra8(rgba)

    def enc(c):
        c = np.nan_to_num(c.astype(np.float64), nan=0.0)
        c = np.clip(c, 0, 1)
        s = np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1 / 2.4) - 0.055)
        return np.floor(s * 255 + 0.5)

    exp_b, exp_g, exp_r = enc(rgba[:, 2]), enc(rgba[:, 1]), enc(rgba[:, 0])
    a = np.clip(np.nan_to_num(rgba[:, 3].astype(np.float64)), 0, 1)
    for got, exp in ((out[:, 0], exp_b), (out[:, 1], exp_g), (out[:, 2], exp_r)):
        assert np.abs(got.astype(int) - exp).max() <= 1  # exact except at rounding ties
        assert (got.astype(int) == exp).mean() > 0.999
    assert np.abs(out[:, 3].astype(int) - np.floor(a * 255 + 0.5)).max() <= 1


def test_generators():
    v, t = scenes.uv_sphere(32, 16, 0.5)
    assert len(t) == 960
    v, t = scenes.icosphere(4)
    assert len(t) == 5120 and len(v) == 2562
    for v, t in (scenes.uv_sphere(32, 16), scenes.icosphere(2), scenes.box((0.1, 0.2, 0.3)), scenes.cylinder(7)):
        p = v["pos"].astype(np.float64)
        g = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
        c = (p[t[:, 0]] + p[t[:, 1]] + p[t[:, 2]]) / 3
        assert ((g * c).sum(1) > 0).all()  # outward winding (convex, centred at 0)
    w = scenes.config2(width=64, height=36)
    assert w.data.n_triangles() == 256 * 5120 + 2 and len(w.data.instances) == 257


def test_scene_validation_errors():
    s = SceneData()
    v, t = scenes.uv_sphere(8, 4)
    g = s.add_geometry(v, t, material())
    s.add_instance(g)
    bad = SceneData()
    t2 = t.copy()
    t2[3, 1] = len(v) + 5
    bad.add_geometry(v, t2, material())
    with pytest.raises(KuafuError, match="vertex index out of range"):
        GpuScene(bad)
    bad = SceneData()
    bad.add_geometry(v, t, 7)  # material 7 does not exist
    with pytest.raises(KuafuError, match="material index out of range"):
        GpuScene(bad)
    bad = SceneData()
    bad.add_geometry(v, t, material())
    bad.add_instance(4)
    with pytest.raises(KuafuError, match="Geometry index is out of bounds"):
        GpuScene(bad)


def test_batched_environment_workload():
    # triangle budgets log-uniform in [lo, hi]: every environment within its
    # range, the instance ranges cover every instance once, one camera each
    w = scenes.config5(n_env=64, size=16, seed=2)
    off = w.data.env_offsets
    assert off[0] == 0 and off[-1] == len(w.data.instances) and (np.diff(off) >= 1).all()
    assert len(w.cam) == 64 and w.height == 16 * 64
    assert (w.env_tris >= 5000 * 0.99).all() and (w.env_tris <= 400000).all()
    # the median of log-uniform [5k, 400k] is sqrt(5k * 400k) ~ 44.7k
    assert 15000 < np.median(w.env_tris) < 130000
    geoms = [g for _, g in w.data.instances]
    tri = np.array([len(w.data.geoms[g][1]) for g in geoms])
    per_env = np.add.reduceat(tri, off[:-1])
    assert (per_env == w.env_tris).all()
    sub = w.data.environment(5)
    assert len(sub.instances) == off[6] - off[5] and sub.env_offsets is None


def test_environment_offset_validation():
    s = SceneData()
    v, t = scenes.uv_sphere(8, 4)
    g = s.add_geometry(v, t, material())
    for _ in range(3):
        s.add_instance(g)
    s.env_offsets = np.array([0, 2, 1, 3], np.uint32)
    with pytest.raises(KuafuError, match="non-decreasing"):
        GpuScene(s)
    s.env_offsets = np.array([0, 1, 2], np.uint32)
    with pytest.raises(KuafuError, match="from 0 to num_instances"):
        GpuScene(s)


def _ktx1(size, faces, internal=0x8C43, big_endian=False, nfaces=6, kvd=b"KTXorientation\x00S=r,T=d\x00\x00"):
    # KTX 1.1 container (Khronos): identifier, 13 u32 header fields, key/value
    # data, then per mip level a u32 image size and the faces (4-byte aligned)
    import struct
    e = ">" if big_endian else "<"
    kv = struct.pack(e + "I", len(kvd)) + kvd
    kv += b"\x00" * ((-len(kv)) % 4)
    hdr = struct.pack(e + "13I", 0x04030201, 0x1401, 1, 0x1908, internal, 0x1908, size, size, 0, 0, nfaces, 1,
                      len(kv))
    body = struct.pack(e + "I", size * size * 4) + b"".join(f.tobytes() for f in faces)
    return b"\xabKTX 11\xbb\r\n\x1a\n" + hdr + kv + body


def _ktx2(size, faces, vk_format=43, scheme=0):
    # KTX 2.0 container: identifier, 9 u32 header fields, the index
    # (DFD/KVD offsets and lengths, SGD u64s), one level index entry, data
    import struct
    data = b"".join(f.tobytes() for f in faces)
    off = 12 + 36 + 32 + 24
    hdr = struct.pack("<9I", vk_format, 1, size, size, 0, 0, 6, 1, scheme)
    index = struct.pack("<4I2Q", 0, 0, 0, 0, 0, 0)
    level = struct.pack("<3Q", off, len(data), len(data))
    return b"\xabKTX 20\xbb\r\n\x1a\n" + hdr + index + level + data


def test_ktx_cubemap_reader(tmp_path):
    import ctypes as C
    from kuafu_amd import lib
    rng = np.random.default_rng(4)
    size = 5
    faces = [rng.integers(0, 256, (size, size, 4), dtype=np.uint8) for _ in range(6)]
    cases = {"a.ktx": _ktx1(size, faces), "b.ktx": _ktx1(size, faces, internal=0x8058, big_endian=True),
             "c.ktx2": _ktx2(size, faces), "d.ktx2": _ktx2(size, faces, vk_format=37)}
    for name, blob in cases.items():
        p = tmp_path / name
        p.write_bytes(blob)
        n = C.c_uint32()
        assert lib().kf_ktx_cubemap_read(str(p).encode(), C.byref(n), None, 0) == 0, name
        assert n.value == size
        out = np.zeros((6, size, size, 4), np.uint8)
        assert lib().kf_ktx_cubemap_read(str(p).encode(), C.byref(n), out.ctypes.data, out.nbytes) == 0
        assert (out == np.stack(faces)).all(), name
    bad = {"five.ktx": (_ktx1(size, faces[:5], nfaces=5), "6 square faces"),
           "bc1.ktx": (_ktx1(size, faces, internal=0x83F1), "uncompressed"),
           "zstd.ktx2": (_ktx2(size, faces, scheme=2), "supercompressed"),
           "cut.ktx": (_ktx1(size, faces)[:-7], "truncated"),
           "png.ktx": (b"\x89PNG\r\n\x1a\n" + b"\x00" * 64, "not a KTX")}
    for name, (blob, msg) in bad.items():
        p = tmp_path / name
        p.write_bytes(blob)
        n = C.c_uint32()
        assert lib().kf_ktx_cubemap_read(str(p).encode(), C.byref(n), None, 0) != 0, name
        assert msg in lib().kf_last_error().decode(), name
