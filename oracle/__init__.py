"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or as the timed CPU baseline.  liboracle.so is
built by oracle/Makefile.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liboracle.so is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, u32, i32, f32 = C.c_void_p, C.c_uint32, C.c_int, C.c_float
        sig = {
            "ko_tea_batch": (None, [vp, vp, u32, vp]),
            "ko_rnd_sequence": (None, [u32, u32, vp, vp]),
            "ko_sincos_batch": (None, [vp, u32, vp, vp]),
            "ko_exp2_batch": (None, [vp, u32, vp]),
            "ko_intersect_triangle": (i32, [vp, vp, f32, f32, vp, vp, vp, vp, vp, vp, vp]),
            "ko_world_to_object": (None, [vp, vp]),
            "ko_scene_create": (vp, [vp]),
            "ko_scene_destroy": (None, [vp]),
            "ko_trace_closest": (None, [vp, vp, u32, vp, i32, i32]),
            "ko_trace_any": (None, [vp, vp, u32, vp, i32, i32]),
            "ko_render": (i32, [vp, vp, vp, u32, u32, u32, u32, u32, i32, i32, vp, vp, vp, vp, vp]),
            "ko_render_region": (i32, [vp, vp, vp, u32, u32, u32, u32, u32, u32, u32, i32, i32, vp, vp, vp, vp,
                                       vp]),
            "ko_relmse": (C.c_double, [vp, vp, u32, u32, u32]),
            "ko_to_bgra8_srgb": (None, [vp, u32, vp]),
            "ko_shading_term": (i32, [i32, i32, vp, u32, vp]),
            "ko_sample": (i32, [i32, i32, u32, vp, u32, vp, vp]),
            "ko_direct": (i32, [i32, vp, vp, vp]),
        }
        for k, (res, args) in sig.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def tea(a, b):
    a = np.ascontiguousarray(a, np.uint32)
    b = np.ascontiguousarray(b, np.uint32)
    out = np.zeros_like(a)
    lib().ko_tea_batch(_p(a), _p(b), a.size, _p(out))
    return out


def rnd_sequence(seed, n):
    out = np.zeros(n, np.float32)
    st = C.c_uint32()
    lib().ko_rnd_sequence(seed, n, _p(out), C.byref(st))
    return out, st.value


def sincos(x):
    x = np.ascontiguousarray(x, np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    lib().ko_sincos_batch(_p(x), x.size, _p(s), _p(c))
    return s, c


def exp2(x):
    x = np.ascontiguousarray(x, np.float32)
    o = np.zeros_like(x)
    lib().ko_exp2_batch(_p(x), x.size, _p(o))
    return o


def intersect_triangle(org, d, tmin, tmax, v0, v1, v2):
    arrs = [np.ascontiguousarray(v, np.float32) for v in (org, d, v0, v1, v2)]  # kept alive for the call
    t, b1, b2 = C.c_float(), C.c_float(), C.c_float()
    back = C.c_uint32()
    h = lib().ko_intersect_triangle(_p(arrs[0]), _p(arrs[1]), tmin, tmax, _p(arrs[2]), _p(arrs[3]), _p(arrs[4]),
                                    C.byref(t), C.byref(b1), C.byref(b2), C.byref(back))
    return (bool(h), t.value, b1.value, b2.value, back.value)


def world_to_object(transform):
    t = np.ascontiguousarray(np.asarray(transform, np.float32).ravel())
    out = np.zeros(12, np.float32)
    lib().ko_world_to_object(_p(t), _p(out))
    return out


class Scene:
    """Oracle copy of a kuafu_amd.api.SceneData (same kf_scene_desc)."""

    def __init__(self, data):
        self._desc = data.desc()
        self.h = lib().ko_scene_create(C.addressof(self._desc))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ko_scene_destroy(self.h)
            self.h = None

    def trace_closest(self, rays, brute=False, threads=0):
        from kuafu_amd import abi
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        hits = np.zeros(len(rays), abi.HIT_DTYPE)
        lib().ko_trace_closest(self.h, _p(rays), len(rays), _p(hits), int(brute), threads)
        return hits

    def trace_any(self, rays, brute=False, threads=0):
        from kuafu_amd import abi
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        lib().ko_trace_any(self.h, _p(rays), len(rays), _p(occ), int(brute), threads)
        return occ

    def render(self, cam, pc, width, height, seed=0, rows=None, precision=32, threads=0, rgba=None,
               keep_samples=False, cols=None):
        """One frame (or the rows / cols ranges of it); returns dict(rgba,
        albedo, normal, samples, counts) of full W x H size."""
        y0, y1 = (0, height) if rows is None else rows
        x0, x1 = (0, width) if cols is None else cols
        rgba = np.zeros((height, width, 4), np.float32) if rgba is None else np.ascontiguousarray(rgba, np.float32)
        alb = np.zeros((height, width, 4), np.float32)
        nrm = np.zeros((height, width, 4), np.float32)
        smp = np.zeros((pc.spp, height, width, 3), np.float32) if keep_samples else None
        cnt = np.zeros(3, np.uint64)
        rc = lib().ko_render_region(self.h, C.addressof(cam), C.addressof(pc), seed, width, height, x0, x1, y0, y1,
                                    precision, threads, _p(rgba), _p(alb), _p(nrm), _p(smp), _p(cnt))
        if rc != 0:
            raise RuntimeError("ko_render failed")
        return dict(rgba=rgba, albedo=alb, normal=nrm, samples=smp, counts=cnt)


def relmse(x, y):
    x = np.ascontiguousarray(x, np.float32).reshape(-1, x.shape[-1])
    y = np.ascontiguousarray(y, np.float32).reshape(-1, y.shape[-1])
    return lib().ko_relmse(_p(x), _p(y), x.shape[0], x.shape[1], y.shape[1])


def to_bgra8(rgba):
    rgba = np.ascontiguousarray(rgba, np.float32)
    out = np.zeros(rgba.shape[:-1] + (4,), np.uint8)
    lib().ko_to_bgra8_srgb(_p(rgba), rgba.size // 4, _p(out))
    return out


# Shading terms and samplers (oracle.h: KO_* ids), for the closed-form pins.
SCHLICK, GGX_D, G1, GEOM_SHADOW, SCHLICK_FRESNEL, REFRACT, REFLECT, LOCAL_TO_WORLD, PERPENDICULAR = range(9)
_TERM_IO = [(2, 1), (2, 1), (2, 1), (3, 1), (2, 1), (7, 3), (6, 3), (6, 3), (3, 3)]
SAMPLE_COSINE, SAMPLE_GGX, SAMPLE_SPHERE, SAMPLE_DISK = range(4)
_SAMPLE_OUT = [3, 3, 3, 2]


def shading_term(term, inputs, precision=64):
    """Evaluate one shading term on rows of inputs (n, k) -> (n, m) float64."""
    k, m = _TERM_IO[term]
    a = np.ascontiguousarray(np.asarray(inputs, np.float64).reshape(-1, k))
    out = np.zeros((len(a), m), np.float64)
    if lib().ko_shading_term(term, precision, _p(a), len(a), _p(out)) != 0:
        raise ValueError("ko_shading_term failed")
    return out[:, 0] if m == 1 else out


def sample(sampler, n, seed=0, param=None, precision=64):
    """n successive samples from one rnd stream; returns (samples, draws)."""
    p = None if param is None else np.ascontiguousarray(np.asarray(param, np.float64).ravel())
    out = np.zeros((n, _SAMPLE_OUT[sampler]), np.float64)
    draws = C.c_uint64()
    if lib().ko_sample(sampler, precision, seed, _p(p), n, _p(out), C.byref(draws)) != 0:
        raise ValueError("ko_sample failed")
    return out, draws.value


def direct(N, V, L, f, a2, diffuse, specular, transmission, inside, emission, seed, precision=64):
    """calcDirectContribution for one shading point; returns (rgb, new seed)."""
    a = np.ascontiguousarray(np.concatenate([np.ravel(N), np.ravel(V), np.ravel(L), [f, a2], np.ravel(diffuse),
                                             np.ravel(specular), np.ravel(transmission), [float(inside)],
                                             np.ravel(emission)]).astype(np.float64))
    assert a.size == 24
    s = C.c_uint32(seed)
    out = np.zeros(3, np.float64)
    if lib().ko_direct(precision, _p(a), C.byref(s), _p(out)) != 0:
        raise ValueError("ko_direct failed")
    return out, s.value
