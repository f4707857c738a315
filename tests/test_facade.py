"""The kuafu:: C++ facade (include/kuafu.hpp) through a C++ client program.

CPU: Config defaults/setters, scene bookkeeping and limits, camera pose and
projection, mesh generators, OBJ import (tests/cpp/facade_test.cpp `cpu`).
GPU: a scene rendered with Kuafu::run() is read back through the API's public
fields only; the same scene rebuilt with the Python host over the C ABI must
give the same frame (the facade's marshalling of geometry, instances,
materials, lights and camera is what is checked).
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_test")


def test_facade_cpu_semantics():
    assert os.path.exists(BIN), "build first (__graft_entry__.build())"
    r = subprocess.run([BIN, "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().endswith("OK")


def c_abi_scene(sc):
    """The dumped facade scene rebuilt with the Python host over the C ABI."""
    from kuafu_amd import GpuScene, abi
    from kuafu_amd.api import SceneData, camera_ubo, material

    s = SceneData()
    for m in sc["materials"]:
        s.add_material(material(diffuse=m["diffuse"], metallic=m["metallic"], specular=m["specular"],
                                roughness=m["roughness"], ior=m["ior"], transmission=m["transmission"],
                                alpha=m["alpha"], emission=m["emission"], emission_strength=m["emissionStrength"]))
    for g in sc["geometries"]:
        v = np.zeros(len(g["pos"]), abi.VERTEX_DTYPE)
        v["pos"] = g["pos"]
        v["normal"] = g["normal"]
        s.add_geometry(v, np.array(g["indices"], np.uint32), g["material"])
    for i in sc["instances"]:
        s.add_instance(i["geometry"], np.array(i["transform"], np.float32).reshape(4, 4))
    d = sc["dir_light"]
    s.set_directional_light(d["direction"], d["color"], d["strength"], d["softness"])
    p = sc["point_light"]
    s.add_point_light(p["position"], p["color"], p["strength"], p["radius"])
    a = sc["active_light"]
    s.add_active_light(np.array(a["view"], np.float32).reshape(4, 4), a["color"], a["fov"], a["strength"])
    c = sc["camera"]
    W, H = c["width"], c["height"]
    cam = camera_ubo(c["position"], c["front"], c["up"], W, H, c["fx"], c["fy"], c["cx"], c["cy"])
    return GpuScene(s), cam, W, H


def render_c_abi(sc):
    from kuafu_amd import Renderer
    from kuafu_amd.api import push_constants
    gs, cam, W, H = c_abi_scene(sc)
    rd = Renderer(W, H)
    for f in range(sc["frames"]):
        pc = push_constants(spp=sc["spp"], path_depth=sc["path_depth"], frame_count=sc.get("first_frame", 0) + f,
                            clear_color=tuple(sc["clear_color"]))
        rd.render(gs, cam, pc, seed=0)
    return rd


@pytest.mark.gpu
def test_facade_render_matches_c_abi(gpu, tmp_path):
    r = subprocess.run([BIN, "gpu", str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    sc = json.load(open(tmp_path / "scene.json"))
    W, H = sc["camera"]["width"], sc["camera"]["height"]
    rgba = np.fromfile(tmp_path / "rgba.bin", np.float32).reshape(H, W, 4)
    bgra = np.fromfile(tmp_path / "bgra.bin", np.uint8).reshape(H, W, 4)
    rd = render_c_abi(sc)
    ours = rd.read(0)
    close = np.abs(ours - rgba) <= 1e-4 * np.maximum(1.0, np.abs(rgba))
    print("facade vs C ABI: %.5f of values within 1e-4" % close.mean())
    assert close.mean() > 0.99
    assert np.abs(ours - rgba).mean() < 1e-3
    assert (np.abs(rd.download_bgra8().astype(int) - bgra.astype(int)) <= 1).mean() > 0.99


@pytest.mark.gpu
def test_facade_incremental_updates_match_fresh_scene(gpu, tmp_path):
    """Kuafu::run() after moved and added instances (TLAS only), a geometry
    deformed in place (BLAS refit) and an edited material, all on the first
    device scene, renders what a scene built from scratch renders."""
    r = subprocess.run([BIN, "gpu_update", str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().endswith("OK")
    sc = json.load(open(tmp_path / "scene.json"))
    W, H = sc["camera"]["width"], sc["camera"]["height"]
    rgba = np.fromfile(tmp_path / "rgba.bin", np.float32).reshape(H, W, 4)
    ours = render_c_abi(sc).read(0)
    close = np.abs(ours - rgba) <= 1e-4 * np.maximum(1.0, np.abs(rgba))
    print("incremental facade vs fresh C ABI scene: %.5f of values within 1e-4" % close.mean())
    assert close.mean() > 0.99 and np.abs(ours - rgba).mean() < 1e-3
