"""ctypes binding of the CPU oracle (liboracle_ptref.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker / baseline.  The product (libpt_hip.so) never
loads the oracle.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_ptref.so"
sys.path.insert(0, str(HERE.parent / "optix-path-tracer_amd"))
import ptamd  # noqa: E402  (struct layouts of include/pt_api.h)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        sigs = {
            "ptref_last_error": (C.c_char_p, []),
            "ptref_tea": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_int32]),
            "ptref_rnd_sequence": (C.c_uint32, [C.c_uint32, C.c_uint32, P]),
            "ptref_scene_create": (P, [C.POINTER(ptamd.SceneDesc)]),
            "ptref_scene_destroy": (None, [P]),
            "ptref_accel_sizes": (None, [P, P]),
            "ptref_accel_read": (None, [P, P, P, P, P]),
            "ptref_trace": (None, [P, P, C.c_uint64, P, P, C.c_int32, C.c_int32]),
            "ptref_occluded": (None, [P, P, C.c_uint64, P]),
            "ptref_render": (C.c_int32, [P, C.POINTER(ptamd.Camera), C.POINTER(ptamd.RenderParams),
                                         C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int32,
                                         P, P, P, P, P]),
            "ptref_display": (None, [P, C.c_uint64, C.c_uint32, P]),
            "ptref_render_pixels": (C.c_int32, [P, C.POINTER(ptamd.Camera), C.POINTER(ptamd.RenderParams),
                                                C.c_uint32, C.c_uint32, P, C.c_uint64, C.c_int32, P, P]),
            "ptref64_last_error": (C.c_char_p, []),
            "ptref64_scene_create": (P, [C.POINTER(ptamd.SceneDesc)]),
            "ptref64_scene_destroy": (None, [P]),
            "ptref64_render_pixels": (C.c_int32, [P, C.POINTER(ptamd.Camera), C.POINTER(ptamd.RenderParams),
                                                  C.c_uint32, C.c_uint32, P, C.c_uint64, C.c_int32, P, P]),
            "ptref_brdf_sample_eval": (None, [C.c_int32, P, P, C.c_float, C.c_float, P, P, P, P, P,
                                              P]),
            "ptref_brdf_eval": (None, [C.c_int32, P, P, C.c_float, C.c_float, P, P, P, P, P]),
            "ptref_cosine_sample": (None, [C.c_float, C.c_float, P, P]),
            "ptref_brdf_sample_many": (None, [C.c_int32, P, P, C.c_float, C.c_float, P, P, C.c_uint32,
                                              C.c_uint64, P, P, P, P]),
            "ptref_metal_scatter_many": (None, [P, P, C.c_float, C.c_uint32, C.c_uint64, P]),
            "ptref_light_sample": (None, [P, C.c_int32, P, P, P, P]),
            "ptref_light_pdf": (C.c_float, [P, C.c_int32, P, P]),
            "ptref_miss": (None, [P, P, P]),
            "ptref_texture": (None, [P, C.c_int32, C.c_float, C.c_float, P, C.c_int32, P]),
            "ptref_hit_record": (None, [P, C.c_int32, C.c_int32, C.c_float, P, P, P, P]),
            "ptref_libm": (None, [C.c_int32, P, P, P, C.c_uint64]),
            "ptref_schlick": (C.c_float, [C.c_float, C.c_float]),
            "ptref_env_tables": (None, [P, P, P, P]),
            "ptref_env_radiance": (None, [P, P, P]),
            "ptref_refract": (C.c_int32, [P, P, C.c_float, P]),
            "ptref_ggx_d": (C.c_float, [P, C.c_float, C.c_float]),
            "ptref_onb": (None, [P, P]),
            "ptref_power_heuristic": (C.c_float, [C.c_float, C.c_float]),
            "ptref_random_points": (C.c_uint32, [C.c_int32, C.c_uint32, C.c_uint32, P]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def p(a):
    # data_as keeps a reference to the array for the lifetime of the pointer
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def f3(v):
    return np.ascontiguousarray(v, np.float32)


def tea(v0, v1, rounds=64):
    return load().ptref_tea(v0, v1, rounds)


def rnd_sequence(seed, n):
    out = np.zeros(n, np.float32)
    st = load().ptref_rnd_sequence(seed, n, p(out))
    return out, st


class OracleScene:
    """The oracle's copy of a scene: float (fp64=False, the GPU's arithmetic)
    or its double build (fp64=True: converged references; only render_pixels
    is available there)."""

    def __init__(self, desc, fp64=False):
        lib = load()
        self.lib = lib
        self.fp64 = fp64
        if fp64:
            self.h = lib.ptref64_scene_create(C.byref(desc))
            err = lib.ptref64_last_error
        else:
            self.h = lib.ptref_scene_create(C.byref(desc))
            err = lib.ptref_last_error
        if not self.h:
            raise RuntimeError("oracle scene: " + err().decode())

    def close(self):
        if self.h:
            (self.lib.ptref64_scene_destroy if self.fp64 else self.lib.ptref_scene_destroy)(self.h)
            self.h = None

    def render_pixels(self, cam, params, frame_begin, frame_count, pix, threads=8, sums=None, stats=None):
        """Sum of de_nan(color()) over frames [frame_begin, +frame_count) for
        each listed pixel (launch index x + W*y): (len(pix), 3) float64."""
        pix = np.ascontiguousarray(pix, np.uint32)
        if sums is None:
            sums = np.zeros((len(pix), 3), np.float64)
        if stats is None:
            stats = np.zeros(4, np.uint64)
        fn = self.lib.ptref64_render_pixels if self.fp64 else self.lib.ptref_render_pixels
        st = fn(self.h, C.byref(cam), C.byref(params), frame_begin, frame_count, p(pix), len(pix), threads,
                p(sums), p(stats))
        if st != 0:
            err = self.lib.ptref64_last_error if self.fp64 else self.lib.ptref_last_error
            raise RuntimeError("oracle render: " + err().decode())
        return sums, stats

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def accel(self):
        sz = np.zeros(3, np.uint64)
        self.lib.ptref_accel_sizes(self.h, p(sz))
        nn, npm, nb = (int(x) for x in sz)
        nodes = np.zeros((nn, 32), np.uint32)
        order = np.zeros(npm, np.uint32)
        morton = np.zeros(npm, np.uint32)
        info = np.zeros((nb, 6), np.uint32)
        self.lib.ptref_accel_read(self.h, p(nodes), p(order), p(morton), p(info))
        return nodes, order, morton, info

    def trace(self, rays, brute=False, counted=False, threads=8):
        rays = np.ascontiguousarray(rays, ptamd.RAY_DTYPE)
        hits = np.zeros(len(rays), ptamd.HIT_DTYPE)
        cnt = np.zeros((len(rays), 3), np.uint32) if counted else None
        self.lib.ptref_trace(self.h, p(rays), len(rays), p(hits), p(cnt), 1 if brute else 0, threads)
        return (hits, cnt) if counted else hits

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, ptamd.RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        self.lib.ptref_occluded(self.h, p(rays), len(rays), p(occ))
        return occ

    def render(self, cam, params, frame_begin, frame_count, threads=8, pix=None, accum=None,
               keep_last=False):
        W, H = params.width, params.height
        if accum is None:
            accum = np.zeros((H, W, 3), np.float64)
        pb, pe = (0, W * H) if pix is None else pix
        rgb = np.zeros((W * H, 3), np.float32) if keep_last else None
        sig = np.zeros(W * H, np.uint32) if keep_last else None
        bounces = np.zeros(W * H, np.uint32) if keep_last else None
        stats = np.zeros(4, np.uint64)
        st = self.lib.ptref_render(self.h, C.byref(cam), C.byref(params), frame_begin, frame_count,
                                   pb, pe, threads, p(accum), p(rgb), p(sig), p(bounces), p(stats))
        if st != 0:
            raise RuntimeError("oracle render: " + self.lib.ptref_last_error().decode())
        return accum, (rgb, sig, bounces), stats

    def light_sample(self, li, P, seed):
        s = C.c_uint32(seed)
        wi = np.zeros(3, np.float32)
        pdf = np.zeros(1, np.float32)
        self.lib.ptref_light_sample(self.h, li, p(f3(P)), C.byref(s), p(wi), p(pdf))
        return wi, float(pdf[0]), s.value

    def light_pdf(self, li, P, wi):
        return float(self.lib.ptref_light_pdf(self.h, li, p(f3(P)), p(f3(wi))))

    def env_tables(self):
        sz = np.zeros(3, np.uint64)
        self.lib.ptref_env_tables(self.h, p(sz), None, None)
        W, H, total = (int(x) for x in sz)
        cond = np.zeros((H, W), np.uint64)
        marg = np.zeros(H, np.uint64)
        if W and H:
            self.lib.ptref_env_tables(self.h, p(sz), p(cond), p(marg))
        return cond, marg, total

    def env_radiance(self, d):
        out = np.zeros(3, np.float32)
        self.lib.ptref_env_radiance(self.h, p(f3(d)), p(out))
        return out

    def miss(self, d):
        out = np.zeros(3, np.float32)
        self.lib.ptref_miss(self.h, p(f3(d)), p(out))
        return out

    def texture(self, tex, u, v, P, index=0):
        out = np.zeros(3, np.float32)
        self.lib.ptref_texture(self.h, tex, u, v, p(f3(P)), index, p(out))
        return out

    def hit_record(self, inst, prim, t, o, d):
        out = np.zeros(11, np.float32)
        mat = C.c_int32()
        self.lib.ptref_hit_record(self.h, inst, prim, t, p(f3(o)), p(f3(d)), p(out), C.byref(mat))
        return out, mat.value


def display(accum, frames):
    H, W, _ = accum.shape
    out = np.zeros((H, W, 4), np.uint8)
    load().ptref_display(p(np.ascontiguousarray(accum)), W * H, frames, p(out))
    return out


def brdf_sample_eval(mtype, c0, c1, p0, p1, wo, n, seed):
    s = C.c_uint32(seed)
    wi = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    load().ptref_brdf_sample_eval(mtype, p(f3(c0)), p(f3(c1)), p0, p1, p(f3(wo)), p(f3(n)), C.byref(s),
                                  p(wi), p(f), p(pdf))
    return wi, f, float(pdf[0]), s.value


def brdf_sample_many(mtype, c0, c1, p0, p1, wo, n, seed0, count):
    """count samples (sample k seeded with tea(k, seed0, 8)): Wi, f, pdf and
    the first rnd() each sample draws."""
    wi = np.zeros((count, 3), np.float32)
    f = np.zeros((count, 3), np.float32)
    pdf = np.zeros(count, np.float32)
    u0 = np.zeros(count, np.float32)
    load().ptref_brdf_sample_many(mtype, p(f3(c0)), p(f3(c1)), p0, p1, p(f3(wo)), p(f3(n)), seed0, count,
                                  p(wi), p(f), p(pdf), p(u0))
    return wi, f, pdf, u0


def metal_scatter_many(wo, n, fuzz, seed0, count):
    d = np.zeros((count, 3), np.float32)
    load().ptref_metal_scatter_many(p(f3(wo)), p(f3(n)), fuzz, seed0, count, p(d))
    return d


def brdf_eval(mtype, c0, c1, p0, p1, wo, wi, n):
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    load().ptref_brdf_eval(mtype, p(f3(c0)), p(f3(c1)), p0, p1, p(f3(wo)), p(f3(wi)), p(f3(n)), p(f),
                           p(pdf))
    return f, float(pdf[0])


LIBM = {"sin": 0, "cos": 1, "tan": 2, "atan": 3, "atan2": 4, "asin": 5, "acos": 6, "pow": 7, "pow5": 8}


def libm(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    yy = None if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), np.float32)
    out = np.zeros_like(x)
    load().ptref_libm(LIBM[fn], p(x), p(yy), p(out), x.size)
    return out


def cosine_sample(u1, u2, n):
    out = np.zeros(3, np.float32)
    load().ptref_cosine_sample(u1, u2, p(f3(n)), p(out))
    return out


def schlick(cosine, ior):
    return float(load().ptref_schlick(cosine, ior))


def refract(v, n, eta):
    out = np.zeros(3, np.float32)
    ok = load().ptref_refract(p(f3(v)), p(f3(n)), eta, p(out))
    return bool(ok), out


def ggx_d(h, nu, nv):
    return float(load().ptref_ggx_d(p(f3(h)), nu, nv))


def onb(n):
    out = np.zeros((3, 3), np.float32)
    load().ptref_onb(p(f3(n)), p(out))
    return out


def power_heuristic(f, g):
    return float(load().ptref_power_heuristic(f, g))


def random_points(kind, seed, n):
    """kind: 0 in unit disk, 1 in unit sphere, 2 on unit sphere."""
    out = np.zeros((n, 3), np.float32)
    st = load().ptref_random_points({"disk": 0, "ball": 1, "sphere": 2}[kind], seed, n, p(out))
    return out, st
