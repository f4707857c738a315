"""Benchmark: path-traced frames of the 1080p icosphere scene on one MI355X.

One step = one frame of the hot path (RayTracer::trace equivalent): 1920x1080,
16 spp, path depth 8, 256 icosphere instances (1.31 M triangles), area + spot
light, gradient environment.  Scene, BVHs and render targets are resident in
HBM before timing starts.

value = rays traced (primary + extension + shadow, counted on the device) /
wall time of the K timed frames / 1e6  [Mrays/s].
roofline = the closest-hit traversal kernel (k_extend): algorithmic bytes per
ray = 32 (ray) + 20 (hit) + 128 * BVH4 node visits + 192 * instance records
(each with its BLAS root node) + 48 * triangle tests, counted by the
instrumented traversal on the same frame, / the kernel's HIP-event time.
cpu_baseline = the FP32 CPU oracle on a bounded band of rows of the same frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s at 1080p; time-to-relMSE 1e-2, denoised vs plain spp, 1 GPU"
SCHEDULE_KEYS = ("samples_per_batch", "refill", "grab", "tail_bounce", "grab0", "grab_shadow", "tail_pairs")


def parse_schedule(text):
    """'spb,refill,grab[,tail_bounce[,grab0[,grab_shadow[,tail_pairs]]]]' -> the
    7 arguments of Renderer.set_schedule (missing ones 0, tail_pairs 1)."""
    v = [int(x) for x in text.split(",")]
    if not 3 <= len(v) <= 7:
        raise SystemExit("bench.py: --schedule takes 3 to 7 comma-separated integers, got %r" % text)
    return (v + [0, 0, 0, 1][len(v) - 3:])[:7]


def host_threads():
    """Threads the CPU leg may use: the process's CPU share (OMP_NUM_THREADS
    where the host sets it, else the CPUs this process may run on)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    return (int(env) if env.isdigit() and int(env) > 0 else aff), aff


def cpu_baseline(workload, budget_s=15.0):
    """Time the FP32 oracle on bands of rows of the same frame (bounded CPU
    sample): 64-row bands from the middle of the frame, wrapping around, until
    about budget_s seconds of rendering have run, on every thread of the
    process's CPU share."""
    import oracle
    threads, aff = host_threads()
    sc = oracle.Scene(workload.data)
    W, H = workload.width, workload.height
    band = min(64, H)
    y = (H // 2) // band * band
    rays, dt, rows = 0, 0.0, 0
    while dt < budget_s:
        y1 = min(y + band, H)
        t = time.perf_counter()
        r = sc.render(workload.cam, workload.pc, W, H, rows=(y, y1), threads=threads)
        dt += time.perf_counter() - t
        rays += int(r["counts"].sum())
        rows += y1 - y
        y = 0 if y1 >= H else y1
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "host_cpu_count": os.cpu_count(), "affinity_cpus": aff,
            "sample": "FP32 CPU oracle (restatement of the reference shaders), %d rows of the %dx%d frame in "
                      "64-row bands from row %d (wrapping to row 0 after the last row), %d spp, %d rays in %.1f s"
                      % (rows, W, H, (H // 2) // band * band, workload.pc.spp, rays, dt)}


def relmse(x, y):
    import numpy as np
    return float(np.mean((x.astype(np.float64) - y) ** 2 / (y.astype(np.float64) ** 2 + 1e-2)))


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16


def unet_flops(height, width, n_images=1):
    """Network FLOPs of one denoise (true channel counts: 9 inputs, 3 outputs;
    the kernels pad to 32 / 16 channels): 2 x 9 x Cin x Cout per pixel and
    3x3 convolution, levels at the padded size (multiple of 16) >> level."""
    ch = (32, 64, 96, 128, 160)
    hp, wp = -(-height // 16) * 16, -(-width // 16) * 16
    px = [(hp >> l) * (wp >> l) for l in range(5)]
    macs, cin = 0, 9
    for l, c in enumerate(ch):  # encoder: cin -> c, c -> c
        macs += px[l] * (cin * c + c * c)
        cin = c
    for l in (3, 2, 1, 0):  # decoder: concat(skip, up) -> c, c -> c
        macs += px[l] * ((ch[l] + ch[l + 1]) * ch[l] + ch[l] * ch[l])
    macs += px[0] * ch[0] * 3  # output layer
    return 2 * 9 * macs * n_images


def unet_rate(den, rd, width, height, n_images, reps=5):
    """Kernel time (HIP events) of one batched denoise of the renderer's
    targets and the MFMA throughput it implies."""
    import numpy as np
    ptr = [rd.device_target(k) for k in range(3)]
    ms = [den.infer_device(ptr[0], ptr[1], ptr[2], width, height, ptr[0], n_images=n_images) for _ in range(reps)]
    k = float(np.median(ms))
    tf = unet_flops(height, width, n_images) / (k * 1e-3) / 1e12
    return {"kernel_ms": round(k, 3), "tflops": round(tf, 1), "frac_of_bf16_peak": round(tf / MFMA_BF16_PEAK_TFLOPS, 3)}


FP64_ROWS = list(range(15, 1080, 30))
FP64_REF = os.path.join(ROOT, "tests", "golden", "fp64_config2_rows_4096spp.npy")


def time_to_relmse(w, gs, rd, target=1e-2, ref_spp=4096):
    """Wall time to reach relMSE <= target against the FP64 CPU oracle: plain
    tracer (spp doubling) vs tracer + denoiser; each point the median of 3
    frames (the same seed: the last frame is the one scored), each frame one
    replay of its captured HIP graph (captured before the timed frames).  The reference
    is the FP64 oracle's 4096-spp render of every 30th row of the 1080p frame
    (36 rows x 1920 px, seed 424242, written by tests/golden/make_fp64_refs.py);
    frames are scored on those rows.  The GPU's own 4096-spp frame is scored
    against it too (the converged-image relMSE)."""
    import numpy as np
    from kuafu_amd.denoiser import Denoiser
    pc = w.pc
    fp64 = os.path.exists(FP64_REF) and (w.width, w.height) == (1920, 1080)
    chunk = 128
    pc.spp = chunk
    for f in range(ref_spp // chunk):
        pc.frame_count = f
        rd.render(gs, w.cam, pc, seed=900000 + f)
    gpu_ref = rd.read(0)[..., :3].astype(np.float64).copy()
    pc.frame_count = 0
    if fp64:
        ref = np.load(FP64_REF).astype(np.float64)
        rows = FP64_ROWS
        score = lambda img: relmse(img[rows], ref)  # noqa: E731
        ref_name = ("FP64 CPU oracle, %d spp, rows 15, 45, ..., 1065 of the frame (36 x 1920 px), seed 424242 "
                    "(tests/golden/fp64_config2_rows_4096spp.npy)" % ref_spp)
    else:
        score = lambda img: relmse(img, gpu_ref)  # noqa: E731
        ref_name = "GPU FP32 tracer, %d spp, same frame (no FP64 rows for this frame size)" % ref_spp

    def timed(fn, reps=3):  # median wall time of reps calls (the last call's result stays)
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    plain = []
    spp = 1
    while spp <= 1024:
        pc.spp = spp
        rd.render(gs, w.cam, pc, seed=31 * spp + 1, use_graph=True)  # capture, untimed
        dt = timed(lambda: rd.render(gs, w.cam, pc, seed=31 * spp + 1, use_graph=True))
        e = score(rd.read(0)[..., :3])
        plain.append({"spp": spp, "ms": round(dt * 1e3, 2), "relmse": round(e, 5)})
        if e <= target:
            break
        spp *= 2
    # the denoised frames are few-sample frames: the schedule is tuned for them
    # (the plain sweep keeps the schedule tuned at 16 spp, whose batches it
    # reaches its target with)
    pc.spp = 1
    tune1 = rd.autotune(gs, w.cam, pc)
    den = Denoiser()
    rd.render(gs, w.cam, pc, seed=5)
    den.apply(rd)  # warm-up: the first call allocates the activations
    dn = []
    spp = 1
    while spp <= 64:
        pc.spp = spp

        def frame():
            rd.render(gs, w.cam, pc, seed=37 * spp + 3, use_graph=True)
            den.apply(rd)

        frame()  # graph capture, untimed
        dt = timed(frame)
        e = score(rd.read(0)[..., :3])
        dn.append({"spp": spp, "ms": round(dt * 1e3, 2), "relmse": round(e, 5)})
        if e <= target:
            break
        spp *= 2
    # denoiser at 1, 2, 4, 8 spp: relMSE of the noisy and the denoised frame
    table = []
    for spp in (1, 2, 4, 8):
        pc.spp = spp
        rd.render(gs, w.cam, pc, seed=41 * spp + 7)
        noisy = score(rd.read(0)[..., :3])
        t = time.perf_counter()
        den.apply(rd)
        dt = time.perf_counter() - t
        table.append({"spp": spp, "noisy_relmse": round(noisy, 5), "denoised_relmse": round(score(rd.read(0)[..., :3]), 5),
                      "denoise_ms": round(dt * 1e3, 2)})
    pc.spp = 16
    hit_p = [p for p in plain if p["relmse"] <= target]
    hit_d = [d for d in dn if d["relmse"] <= target]
    res = {"target_relmse": target, "reference": ref_name,
           "gpu_%dspp_relmse" % ref_spp: round(score(gpu_ref), 6) if fp64 else None,
           "plain": plain, "denoised": dn, "denoiser_by_spp": table,
           "unet": unet_rate(den, rd, w.width, w.height, 1),
           "denoised_schedule": {k: tune1[k] for k in SCHEDULE_KEYS if k in tune1}}
    if hit_p and hit_d:
        res["speedup"] = round(hit_p[0]["ms"] / hit_d[0]["ms"], 2)
    return res


def denoised_frames(w, gs, rd, frames, warmup, schedule=""):
    """The denoised point of the time-to-relMSE sweep on its own: 1-spp frames
    of the 1080p scene (each a replay of the frame's HIP graph), each followed
    by the denoise, with the schedule tuned for 1-spp frames (or the given
    one).  Host wall time per frame."""
    import numpy as np
    from kuafu_amd.denoiser import Denoiser
    pc = w.pc
    pc.spp = 1
    pc.frame_count = 0
    tune = None
    if schedule:
        rd.set_schedule(*parse_schedule(schedule))
    else:
        tune = rd.autotune(gs, w.cam, pc)
    den = Denoiser()
    wall, trace = [], []
    for k in range(warmup + frames):
        t = time.perf_counter()
        st = rd.render(gs, w.cam, pc, seed=37 + k, use_graph=True)
        den.apply(rd)
        if k >= warmup:
            wall.append((time.perf_counter() - t) * 1e3)
            trace.append(st.trace_ms)
    return {"workload": "config2_icospheres256_1080p_1spp_denoised", "frames": frames,
            "frame_ms_p50": round(float(np.median(wall)), 3), "frame_ms_min": round(min(wall), 3),
            "trace_ms_p50": round(float(np.median(trace)), 3),
            "schedule": None if tune is None else {k: tune[k] for k in SCHEDULE_KEYS if k in tune}}


def batched_environments(frames=8, schedule=None, full_tlas_camera=False):
    """128 environments of 512x512 (8 spp, depth 6, 5k-400k triangles each):
    autotuned schedule, every frame one replay of a captured HIP graph, then
    one batched denoise of the 128 images.  Frame latency = render + denoise,
    host wall time per frame."""
    import numpy as np
    from kuafu_amd import GpuScene, Renderer, scenes
    from kuafu_amd.denoiser import Denoiser
    w = scenes.config5()
    gs = GpuScene(w.data)
    info = gs.bvh_info()
    rd = Renderer(w.width, w.height)
    pc = w.pc
    pc.frame_count = 0
    if schedule:
        rd.set_schedule(*schedule)
        tune = dict(zip(SCHEDULE_KEYS, schedule), default_ms=0.0, frame_ms=0.0)
    else:
        tune = rd.autotune(gs, w.cam, pc)
    den = Denoiser()
    rd.render_envs(gs, w.cam, pc, seed=1, use_graph=True, full_tlas_camera=full_tlas_camera)  # capture
    den.apply(rd)
    lat, rend, dms = [], [], []
    rays = 0
    kms = 0.0
    for k in range(frames):
        t0 = time.perf_counter()
        st = rd.render_envs(gs, w.cam, pc, seed=2 + k, use_graph=True, full_tlas_camera=full_tlas_camera)
        t1 = time.perf_counter()
        den.apply(rd)
        t2 = time.perf_counter()
        lat.append((t2 - t0) * 1e3)
        rend.append((t1 - t0) * 1e3)
        dms.append((t2 - t1) * 1e3)
        rays += st.rays()
        kms += st.trace_ms
    res = {"environments": w.n_env, "width": w.env_height, "height": w.env_height, "spp": pc.spp,
           "path_depth": pc.path_depth, "triangles_instanced": int(w.env_tris.sum()),
           "triangles_per_env_min_max": [int(w.env_tris.min()), int(w.env_tris.max())],
           "unique_triangles": int(info.triangles), "instances": int(info.instances),
           "bvh_build_ms": round(float(info.build_ms), 2), "rays_per_frame": rays // frames,
           "mrays_s": round(rays / (sum(rend) * 1e-3) / 1e6, 1),
           "kernel_mrays_s": round(rays / (kms * 1e-3) / 1e6, 1),
           "frame_ms_p50": round(float(np.percentile(lat, 50)), 2),
           "frame_ms_p99": round(float(np.percentile(lat, 99)), 2),
           "render_ms_p50": round(float(np.percentile(rend, 50)), 2),
           "denoise_ms_p50": round(float(np.percentile(dms, 50)), 2), "frames": frames,
           "schedule": {k: tune[k] for k in SCHEDULE_KEYS if k in tune},
           "autotune_ms": [round(tune["default_ms"], 1), round(tune["frame_ms"], 1)],
           "unet": unet_rate(den, rd, w.width, w.env_height, w.n_env)}
    gs.close()
    rd.close()
    return res


def active_light_frame(steps=3):
    """The active-light sensing frame (1280x720, 32 spp, depth 6): left and
    right IR cameras (55 mm baseline) and the RGB camera rendered one after
    another, as setCamera + run + download per camera; the IR images come back
    as FP32 single-channel images, the RGB image as BGRA8 (downloadLatestFrame).
    Host wall time per three-camera frame, downloads included."""
    import numpy as np
    from kuafu_amd import GpuScene, Renderer, scenes
    ws = [scenes.config3(camera=c) for c in ("left", "right", "rgb")]
    w = ws[0]
    gs = GpuScene(w.data)
    rd = Renderer(w.width, w.height)
    pc = w.pc
    pc.frame_count = 0
    tune = rd.autotune(gs, w.cam, pc)

    def frame(seed):
        rays, kms = 0, 0.0
        for k, cw in enumerate(ws):
            st = rd.render(gs, cw.cam, pc, seed=seed + k)
            rays += st.rays()
            kms += st.trace_ms
            if k < 2:
                rd.download_gray32f()
            else:
                rd.download_bgra8()
        return rays, kms

    frame(1)
    lat, rays, kms = [], 0, 0.0
    for k in range(steps):
        t = time.perf_counter()
        r, m = frame(10 + 3 * k)
        lat.append((time.perf_counter() - t) * 1e3)
        rays += r
        kms += m
    gs.close()
    rd.close()
    return {"workload": "config3_active_light_720p_32spp_3cameras", "cameras": ["ir_left", "ir_right", "rgb"],
            "outputs": ["f32 single channel", "f32 single channel", "bgra8 srgb"],
            "width": w.width, "height": w.height, "spp": pc.spp, "path_depth": pc.path_depth,
            "frame_ms_p50": round(float(np.median(lat)), 2), "rays_per_frame": rays // steps,
            "mrays_s": round(rays / (sum(lat) * 1e-3) / 1e6, 1),
            "kernel_mrays_s": round(rays / (kms * 1e-3) / 1e6, 1),
            "schedule": {k: tune[k] for k in SCHEDULE_KEYS if k in tune}}


def other_workloads(steps=3):
    """Mrays/s of the other scene recipes on the same build (not the headline)."""
    from kuafu_amd import GpuScene, Renderer, scenes
    out = {}
    for w in (scenes.config1(),):
        gs = GpuScene(w.data)
        rd = Renderer(w.width, w.height)
        pc = w.pc
        pc.frame_count = 0
        rd.autotune(gs, w.cam, pc)
        rd.render(gs, w.cam, pc, seed=1)
        rays = 0
        t = time.perf_counter()
        for k in range(steps):
            rays += rd.render(gs, w.cam, pc, seed=2 + k).rays()
        dt = time.perf_counter() - t
        out[w.name] = {"width": w.width, "height": w.height, "spp": pc.spp, "path_depth": pc.path_depth,
                       "mrays_s": round(rays / dt / 1e6, 1), "ms_per_frame": round(dt / steps * 1e3, 2)}
        gs.close()
        rd.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--samples-per-batch", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-relmse", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--schedule", default="",
                    help="samples_per_batch,refill,grab[,tail_bounce[,grab0[,grab_shadow[,tail_pairs]]]] instead "
                         "of autotuning")
    ap.add_argument("--envs-only", action="store_true",
                    help="only the 128-environment workload (frames = --steps), e.g. for profiling")
    ap.add_argument("--denoised-only", action="store_true",
                    help="only 1-spp frames + denoise of the 1080p scene (frames = --steps), e.g. for profiling")
    ap.add_argument("--separate-geometry", action="store_true",
                    help="every geometry traverses its own BVH, also where geometries have the same triangles "
                         "(comparison)")
    ap.add_argument("--full-tlas-camera", action="store_true",
                    help="camera rays start from the whole TLAS instead of their pixel tile's root (comparison)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU; --gpus %d is not supported (rendering would shard by image "
              "rows or environments with the scene replicated, denoiser training by data parallelism)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from kuafu_amd import GpuScene, Renderer, lib, scenes
    from kuafu_amd.abi import check

    check(lib().kf_init(0))
    if args.envs_only:
        sched = parse_schedule(args.schedule) if args.schedule else None
        print(json.dumps(batched_environments(frames=args.steps, schedule=sched,
                                              full_tlas_camera=args.full_tlas_camera)))
        return
    w = scenes.config2(width=args.width, height=args.height)
    w.data.separate_geometry = args.separate_geometry
    print("bench: building scene (%d triangles)" % w.data.n_triangles(), file=sys.stderr, flush=True)
    gs = GpuScene(w.data)
    info = gs.bvh_info()
    rd = Renderer(w.width, w.height)
    pc = w.pc
    pc.frame_count = 0

    if args.denoised_only:
        print(json.dumps(denoised_frames(w, gs, rd, args.steps, args.warmup, args.schedule)))
        return

    # schedule (batch size, refill threshold, reservation per wave) picked by
    # timing frames of this scene, before anything is measured
    tune = None
    if args.schedule:
        v = parse_schedule(args.schedule)
        rd.set_schedule(*v)
        tune = dict(zip(SCHEDULE_KEYS, v))
    elif not args.no_autotune and not args.samples_per_batch:
        print("bench: autotune", file=sys.stderr, flush=True)
        tune = rd.autotune(gs, w.cam, pc)
    # instrumented frame (untimed): node visits / triangle tests of k_extend
    ft = args.full_tlas_camera
    cst = rd.render(gs, w.cam, pc, seed=1, samples_per_batch=args.samples_per_batch, count_traversal=True,
                    full_tlas_camera=ft)
    for _ in range(args.warmup):
        rd.render(gs, w.cam, pc, seed=0, samples_per_batch=args.samples_per_batch, full_tlas_camera=ft)
    print("bench: warmup done", file=sys.stderr, flush=True)
    rays = 0
    ext_rays = 0
    ext_ms = 0.0
    trace_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = rd.render(gs, w.cam, pc, seed=100 + k, samples_per_batch=args.samples_per_batch, full_tlas_camera=ft)
        rays += st.rays()
        # rays k_extend traced (the tail kernel's are not in extend_ms)
        ext_rays += st.primary_rays + st.extension_rays + st.speculative_rays - st.tail_rays
        ext_ms += st.extend_ms
        trace_ms += st.trace_ms
        launches += st.extend_launches
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = rays / wall / 1e6

    closest = cst.primary_rays + cst.extension_rays + cst.speculative_rays  # rays k_extend traced
    nodes_per_ray = cst.node_visits / max(closest, 1)
    insts_per_ray = cst.instance_visits / max(closest, 1)
    tris_per_ray = cst.tri_tests / max(closest, 1)
    bytes_per_ray = 32 + 20 + 128 * nodes_per_ray + 192 * insts_per_ray + 48 * tris_per_ray
    achieved = bytes_per_ray * ext_rays / (ext_ms * 1e-3) / 1e9
    # HBM traffic per launch is not measurable from inside the process: it
    # comes from the committed rocprofv3 counter profile of this kernel on
    # this scene and schedule (FETCH_SIZE and WRITE_SIZE, each scaled by what
    # the counter reads for a copy of known size), named with its build.
    traffic, tsrc = None, None
    tpath = os.path.join(ROOT, "profiles", "extend_traffic_config2.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_launch")
        tsrc = ("profiles/extend_traffic_config2.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of k_extend<false>, "
                "schedule %s, build %s, %s)" % (tj.get("schedule", "?"), tj.get("build", "?"), tj.get("date", "?")))
    avg_ms = ext_ms / max(launches, 1)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
            "dram_gbs": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
            "dram_frac": round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
            "kernel": "k_extend (closest-hit traversal)",
            "bytes_per_ray": round(bytes_per_ray, 1), "node_visits_per_ray": round(nodes_per_ray, 2),
            "instance_visits_per_ray": round(insts_per_ray, 2), "tri_tests_per_ray": round(tris_per_ray, 2),
            "tlas_node_visits_per_ray": round(cst.tlas_node_visits / max(closest, 1), 2),
            "avg_launch_ms": round(avg_ms, 4)}
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "width": w.width, "height": w.height, "spp": pc.spp,
                   "path_depth": pc.path_depth, "triangles": int(info.triangles), "instances": int(info.instances),
                   "rays_per_frame": rays // args.steps,
                   "schedule": None if tune is None else {k: tune[k] for k in SCHEDULE_KEYS if k in tune},
                   "autotune_ms": None if tune is None or "frame_ms" not in tune else
                   [round(tune["default_ms"], 2), round(tune["frame_ms"], 2)],
                   "kernel_mrays_s": round(rays / (trace_ms * 1e-3) / 1e6, 2)},
        "roofline": roof,
    }
    import ctypes
    bw = ctypes.c_float()
    check(lib().kf_bench_copy(1 << 30, 10, ctypes.byref(bw)))
    roof["achievable_copy_gbs"] = round(bw.value, 1)
    roof["frac_of_achievable"] = round(achieved / bw.value, 4)
    if not args.no_relmse:
        print("bench: time to relMSE", file=sys.stderr, flush=True)
        out["time_to_relmse"] = time_to_relmse(w, gs, rd)
        print("bench: other scenes", file=sys.stderr, flush=True)
        out["other_workloads"] = other_workloads()
        out["other_workloads"]["config3_active_light_720p_32spp_3cameras"] = active_light_frame()
        print("bench: batched environments", file=sys.stderr, flush=True)
        out["other_workloads"]["config5_envs128_512x512_8spp"] = batched_environments()
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
