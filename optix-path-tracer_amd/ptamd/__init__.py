"""ctypes binding of libpt_hip.so (include/pt_api.h, include/pt_scene.h).

The library is the MI355X path tracer: the scene, acceleration structure,
render and query entry points of the C ABI.  This module only loads it and
mirrors its structs; it has no compute of its own and no fallback: if the
library is missing, loading fails with an error.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
REPO = PKG_DIR.parent
LIB_PATH = PKG_DIR / "libpt_hip.so"

PT_OK = 0

PT_AXIS_X, PT_AXIS_Y, PT_AXIS_Z = 0, 1, 2
PT_GEOM_SPHERES, PT_GEOM_RECTS, PT_GEOM_BOXES, PT_GEOM_MESH, PT_GEOM_VOLUMES = 0, 1, 2, 3, 4
PT_VOLUME_SPHERE, PT_VOLUME_BOX = 0, 1
PT_GEOM_CYLINDERS, PT_GEOM_MOVING_SPHERES = 5, 6
PT_MAT_LAMBERTIAN, PT_MAT_METAL, PT_MAT_DIELECTRIC, PT_MAT_DIFFUSE_LIGHT = 0, 1, 2, 3
PT_MAT_ASHIKHMIN_SHIRLEY, PT_MAT_TORRANCE_SPARROW, PT_MAT_OREN_NAYAR = 4, 5, 6
PT_MAT_ISOTROPIC, PT_MAT_NORMAL = 7, 8
PT_TEX_CONSTANT, PT_TEX_IMAGE, PT_TEX_CHECKER, PT_TEX_NOISE, PT_TEX_GRADIENT = 0, 1, 2, 3, 4
PT_NOISE_FLOATS = 1792
PT_TEX_VECTOR = 5
PT_IMAGE_PNG, PT_IMAGE_HDR = 0, 1
PT_LIGHT_RECT, PT_LIGHT_SPHERE, PT_LIGHT_ENV = 0, 1, 2
PT_MISS_GRADIENT, PT_MISS_CONSTANT, PT_MISS_IMAGE, PT_MISS_HDR = 0, 1, 2, 3
PT_FLAG_MATERIAL_SORT = 0x1
PT_FLAG_KEEP_PATHS = 0x2
PT_FLAG_ENV_MIS = 0x4
PT_FLAG_COUNT_TRAVERSAL = 0x8
PT_FLAG_RADIANCE_CACHE = 0x10
PT_FLAG_GRAPH = 0x20
PT_FLAG_TUNED = 0x40
PT_FLAG_DIRECT = 0x80
PT_SCENE_RANDOM_SPHERES, PT_SCENE_CORNELL, PT_SCENE_MESHES, PT_SCENE_BATCH, PT_SCENE_MOVING_SPHERES = 0, 1, 2, 3, 4
PT_SCENE_NEXT_WEEK = 5
PT_CACHE_INPUTS, PT_CACHE_OUTPUTS, PT_CACHE_PARAMS = 96, 3, 18883


class SphereDesc(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float), ("material", C.c_int32)]


class RectDesc(C.Structure):
    _fields_ = [("axis", C.c_int32), ("flip", C.c_int32), ("a0", C.c_float), ("a1", C.c_float),
                ("b0", C.c_float), ("b1", C.c_float), ("k", C.c_float), ("material", C.c_int32)]


class BoxDesc(C.Structure):
    _fields_ = [("pmin", C.c_float * 3), ("pmax", C.c_float * 3), ("material", C.c_int32)]


class VolumeDesc(C.Structure):
    _fields_ = [("shape", C.c_int32), ("a", C.c_float * 3), ("b", C.c_float * 3), ("density", C.c_float),
                ("material", C.c_int32)]


class CylinderDesc(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("length", C.c_float), ("radius", C.c_float), ("material", C.c_int32)]


class MovingSphereDesc(C.Structure):
    _fields_ = [("center0", C.c_float * 3), ("center1", C.c_float * 3), ("radius", C.c_float),
                ("time0", C.c_float), ("time1", C.c_float), ("material", C.c_int32)]


class MeshDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("vertices", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
                ("num_triangles", C.c_uint32), ("indices", C.POINTER(C.c_uint32)),
                ("face_materials", C.POINTER(C.c_int32)), ("material", C.c_int32),
                ("face_indices", C.POINTER(C.c_int32))]


class InstanceDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("first", C.c_uint32), ("count", C.c_uint32),
                ("has_transform", C.c_int32), ("obj_to_world", C.c_float * 12),
                ("world_to_obj", C.c_float * 12)]


class MaterialDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("tex", C.c_int32 * 2), ("param", C.c_float * 4)]


class TextureDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("rgb", C.c_float * 3), ("width", C.c_uint32),
                ("height", C.c_uint32), ("texels", C.POINTER(C.c_float))]


class LightDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("axis", C.c_int32), ("a0", C.c_float), ("a1", C.c_float),
                ("b0", C.c_float), ("b1", C.c_float), ("k", C.c_float), ("center", C.c_float * 3),
                ("radius", C.c_float), ("emission", C.c_float * 3)]


class MissDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("color1", C.c_float * 3), ("color2", C.c_float * 3),
                ("texture", C.c_int32), ("spherical", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [("num_spheres", C.c_uint32), ("spheres", C.POINTER(SphereDesc)),
                ("num_rects", C.c_uint32), ("rects", C.POINTER(RectDesc)),
                ("num_boxes", C.c_uint32), ("boxes", C.POINTER(BoxDesc)),
                ("num_meshes", C.c_uint32), ("meshes", C.POINTER(MeshDesc)),
                ("num_instances", C.c_uint32), ("instances", C.POINTER(InstanceDesc)),
                ("num_materials", C.c_uint32), ("materials", C.POINTER(MaterialDesc)),
                ("num_textures", C.c_uint32), ("textures", C.POINTER(TextureDesc)),
                ("num_lights", C.c_uint32), ("lights", C.POINTER(LightDesc)),
                ("miss", MissDesc),
                ("num_volumes", C.c_uint32), ("volumes", C.POINTER(VolumeDesc)),
                ("num_cylinders", C.c_uint32), ("cylinders", C.POINTER(CylinderDesc)),
                ("num_moving_spheres", C.c_uint32), ("moving_spheres", C.POINTER(MovingSphereDesc))]


class Camera(C.Structure):
    _fields_ = [("lower_left", C.c_float * 3), ("horizontal", C.c_float * 3),
                ("vertical", C.c_float * 3), ("origin", C.c_float * 3), ("u", C.c_float * 3),
                ("v", C.c_float * 3), ("w", C.c_float * 3), ("lens_radius", C.c_float),
                ("time0", C.c_float), ("time1", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_uint32),
                ("russian", C.c_int32), ("rr_min_depth", C.c_uint32),
                ("frames_per_batch", C.c_uint32), ("flags", C.c_uint32), ("reserved", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("node_fetches", C.c_uint64), ("prim_fetches", C.c_uint64),
                ("instance_fetches", C.c_uint64), ("shadow_node_fetches", C.c_uint64),
                ("shadow_prim_fetches", C.c_uint64), ("shadow_instance_fetches", C.c_uint64),
                ("exhausted_paths", C.c_uint64), ("cache_queries", C.c_uint64),
                ("cache_train_vertices", C.c_uint64), ("cache_loss", C.c_double), ("cache_steps", C.c_uint32), ("gpu_ms", C.c_double), ("extend_ms", C.c_double),
                ("shade_ms", C.c_double), ("extend_launches", C.c_uint32),
                ("shade_launches", C.c_uint32), ("nan_paths", C.c_uint64),
                ("cache_judge", C.c_double), ("cache_fallbacks", C.c_uint32),
                ("cache_train_ms", C.c_double), ("cache_infer_ms", C.c_double),
                ("cache_train_samples", C.c_uint64), ("cache_infer_samples", C.c_uint64),
                ("cache_plain_batches", C.c_uint32)]


class AccelStats(C.Structure):
    _fields_ = [("num_blas", C.c_uint32), ("num_instances", C.c_uint32),
                ("num_primitives", C.c_uint64), ("num_nodes", C.c_uint64), ("bytes", C.c_uint64),
                ("build_ms", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("frames_per_batch", C.c_uint32), ("material_sort", C.c_int32), ("graph", C.c_int32),
                ("candidates", C.c_uint32), ("ms_per_frame", C.c_double), ("paths_per_batch", C.c_uint64)]


class SceneInfo(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_depth", C.c_uint32), ("russian", C.c_int32), ("rr_min_depth", C.c_uint32),
                ("flags", C.c_uint32), ("reserved", C.c_uint32)]


RAY_DTYPE = np.dtype([("org", np.float32, 3), ("tmin", np.float32), ("dir", np.float32, 3),
                      ("tmax", np.float32)])
HIT_DTYPE = np.dtype([("t", np.float32), ("inst", np.int32), ("prim", np.int32),
                      ("reserved", np.uint32)])

_lib = None


def _declare(lib):
    P = C.c_void_p
    sigs = {
        "pt_api_version": (C.c_int32, []),
        "pt_create": (C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
        "pt_destroy": (None, [P]),
        "pt_last_error": (C.c_char_p, [P]),
        "pt_set_scene": (C.c_int32, [P, C.POINTER(SceneDesc)]),
        "pt_build_accel": (C.c_int32, [P, C.POINTER(AccelStats)]),
        "pt_set_camera": (C.c_int32, [P, C.POINTER(Camera)]),
        "pt_render": (C.c_int32, [P, C.POINTER(RenderParams), C.c_uint32, C.c_uint32,
                                  C.POINTER(RenderStats)]),
        "pt_render_batch": (C.c_int32, [C.POINTER(C.c_void_p), C.POINTER(RenderParams), C.POINTER(C.c_uint32),
                                        C.c_uint32, C.POINTER(RenderStats)]),
        "pt_sync": (C.c_int32, [P]),
        "pt_save_image": (C.c_int32, [P, C.c_char_p, C.c_int32]),
        "pt_read_accum": (C.c_int32, [P, P, C.POINTER(C.c_uint32)]),
        "pt_read_display": (C.c_int32, [P, P]),
        "pt_read_paths": (C.c_int32, [P, P, P, P]),
        "pt_trace_closest": (C.c_int32, [P, P, C.c_uint64, P]),
        "pt_trace_closest_device": (C.c_int32, [P, P, C.c_uint64, P]),
        "pt_trace_closest_device_noalloc": (C.c_int32, [P, P, C.c_uint64, P, P]),
        "pt_trace_closest_counted": (C.c_int32, [P, P, C.c_uint64, P, P]),
        "pt_trace_occluded": (C.c_int32, [P, P, C.c_uint64, P]),
        "pt_read_accel": (C.c_int32, [P, P, C.POINTER(C.c_uint64), P, P, C.POINTER(C.c_uint64), P,
                                      C.POINTER(C.c_uint32)]),
        "pt_read_env_tables": (C.c_int32, [P, P, P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint64)]),
        "pt_autotune": (C.c_int32, [P, C.POINTER(RenderParams), C.c_uint32, C.POINTER(TuneResult)]),
        "pt_autotune_batch": (C.c_int32, [C.POINTER(C.c_void_p), C.POINTER(RenderParams), C.POINTER(C.c_uint32),
                                          C.c_uint32, C.POINTER(TuneResult)]),
        "pt_cache_init": (C.c_int32, [P, C.c_uint32]),
        "pt_cache_configure": (C.c_int32, [P, C.c_uint32, C.c_float, C.c_uint32, C.c_uint32, C.c_float]),
        "pt_cache_judge_configure": (C.c_int32, [P, C.c_float, C.c_uint32]),
        "pt_cache_read_records": (C.c_int32, [P, C.c_int32, C.c_uint32, P, P, P, C.POINTER(C.c_uint32)]),
        "pt_cache_params": (C.c_int32, [P, P, P]),
        "pt_cache_infer": (C.c_int32, [P, P, C.c_uint64, P]),
        "pt_cache_train": (C.c_int32, [P, P, P, C.c_uint64, C.c_float, C.POINTER(C.c_float), P]),
        "pt_scene_build": (C.c_int32, [C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_void_p)]),
        "pt_scene_build_obj": (C.c_int32, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32, C.c_uint32,
                                           C.POINTER(C.c_void_p)]),
        "pt_scene_get_desc": (C.POINTER(SceneDesc), [P]),
        "pt_scene_get_camera": (C.POINTER(Camera), [P]),
        "pt_scene_get_info": (C.POINTER(SceneInfo), [P]),
        "pt_scene_free": (None, [P]),
        "pt_scene_last_error": (C.c_char_p, []),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def load():
    """Load libpt_hip.so (built in-tree); raises if it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        _lib = _declare(C.CDLL(str(LIB_PATH)))
    return _lib


def ptr(a: np.ndarray):
    # data_as keeps a reference to the array for the lifetime of the pointer
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class PtError(RuntimeError):
    pass


class Scene:
    """A synthetic scene built by the host scene API inside the library."""

    def __init__(self, config: int, width: int = 0, height: int = 0, variant: int = 0, _handle=None):
        lib = load()
        h = _handle
        if h is None:
            h = C.c_void_p()
            st = lib.pt_scene_build(config, width, height, variant, C.byref(h))
            if st != PT_OK:
                raise PtError(f"pt_scene_build: {lib.pt_scene_last_error().decode()}")
        self._lib = lib
        self.handle = h
        self.desc = lib.pt_scene_get_desc(h).contents
        self.camera = lib.pt_scene_get_camera(h).contents
        self.info = lib.pt_scene_get_info(h).contents

    @classmethod
    def from_obj(cls, obj_file, assets_folder=None, hdr_file=None, width=0, height=0):
        """pt_scene_build_obj: an OBJ/MTL model under an .hdr environment or
        the gradient sky."""
        lib = load()
        h = C.c_void_p()
        enc = lambda x: None if x is None else str(x).encode()
        st = lib.pt_scene_build_obj(enc(obj_file), enc(assets_folder), enc(hdr_file), width, height, C.byref(h))
        if st != PT_OK:
            raise PtError(f"pt_scene_build_obj: {lib.pt_scene_last_error().decode()}")
        return cls(-1, _handle=h)

    def close(self):
        if self.handle:
            self._lib.pt_scene_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render_params(self, flags=None, frames_per_batch=0):
        """Render parameters of the workload; `flags` defaults to the
        scene's own PT_FLAG_* set."""
        i = self.info
        return RenderParams(i.width, i.height, i.max_depth, i.russian, i.rr_min_depth,
                            frames_per_batch, i.flags if flags is None else flags, 0)


class Context:
    """One pt_context on a GPU."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.pt_create(device, C.byref(h))
        if st != PT_OK:
            raise PtError(f"pt_create({device}): {self.lib.pt_last_error(None).decode()}")
        self.h = h

    def _check(self, st, what):
        if st != PT_OK:
            raise PtError(f"{what} -> {st}: {self.lib.pt_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.pt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_scene(self, desc: SceneDesc):
        self._check(self.lib.pt_set_scene(self.h, C.byref(desc)), "pt_set_scene")

    def build_accel(self) -> AccelStats:
        s = AccelStats()
        self._check(self.lib.pt_build_accel(self.h, C.byref(s)), "pt_build_accel")
        return s

    def set_camera(self, cam: Camera):
        self._check(self.lib.pt_set_camera(self.h, C.byref(cam)), "pt_set_camera")

    def render(self, params: RenderParams, frame_begin: int, frame_count: int, stats=True):
        s = RenderStats() if stats else None
        self._check(self.lib.pt_render(self.h, C.byref(params), frame_begin, frame_count,
                                       C.byref(s) if s is not None else None), "pt_render")
        return s

    def sync(self):
        self._check(self.lib.pt_sync(self.h), "pt_sync")

    def read_accum(self, w: int, h: int):
        a = np.zeros((h, w, 3), np.float64)
        fr = C.c_uint32()
        self._check(self.lib.pt_read_accum(self.h, ptr(a), C.byref(fr)), "pt_read_accum")
        return a, fr.value

    def save_image(self, path, fmt=None):
        """pt_save_image: PNG (sqrt gamma) or Radiance HDR (c / (1 + c)) of
        the accumulation average; the format follows the suffix unless given."""
        if fmt is None:
            fmt = PT_IMAGE_HDR if str(path).lower().endswith(".hdr") else PT_IMAGE_PNG
        self._check(self.lib.pt_save_image(self.h, str(path).encode(), fmt), "pt_save_image")

    def read_display(self, w: int, h: int):
        a = np.zeros((h, w, 4), np.uint8)
        self._check(self.lib.pt_read_display(self.h, ptr(a)), "pt_read_display")
        return a

    def read_paths(self, w: int, h: int):
        rgb = np.zeros((w * h, 3), np.float32)
        sig = np.zeros(w * h, np.uint32)
        b = np.zeros(w * h, np.uint32)
        self._check(self.lib.pt_read_paths(self.h, ptr(rgb), ptr(sig), ptr(b)), "pt_read_paths")
        return rgb, sig, b

    def trace_closest(self, rays: np.ndarray, counted=False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        if counted:
            cnt = np.zeros((len(rays), 3), np.uint32)
            self._check(self.lib.pt_trace_closest_counted(self.h, ptr(rays), len(rays), ptr(hits),
                                                          ptr(cnt)), "pt_trace_closest_counted")
            return hits, cnt
        self._check(self.lib.pt_trace_closest(self.h, ptr(rays), len(rays), ptr(hits)),
                    "pt_trace_closest")
        return hits

    def trace_occluded(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        self._check(self.lib.pt_trace_occluded(self.h, ptr(rays), len(rays), ptr(occ)),
                    "pt_trace_occluded")
        return occ

    def read_env_tables(self):
        """Environment importance-sampling tables: (cond (H, W) uint64,
        marg (H,) uint64, total)."""
        w, h, tot = C.c_uint32(), C.c_uint32(), C.c_uint64()
        self._check(self.lib.pt_read_env_tables(self.h, None, None, C.byref(w), C.byref(h), C.byref(tot)),
                    "pt_read_env_tables")
        cond = np.zeros((h.value, w.value), np.uint64)
        marg = np.zeros(h.value, np.uint64)
        if w.value:
            self._check(self.lib.pt_read_env_tables(self.h, ptr(cond), ptr(marg), C.byref(w), C.byref(h),
                                                    C.byref(tot)), "pt_read_env_tables")
        return cond, marg, tot.value

    def autotune(self, params, frames):
        r = TuneResult()
        self._check(self.lib.pt_autotune(self.h, C.byref(params), frames, C.byref(r)), "pt_autotune")
        return r

    # ------------------------------------------------ radiance cache
    def cache_init(self, seed=0):
        self._check(self.lib.pt_cache_init(self.h, seed), "pt_cache_init")

    def cache_params(self, new=None):
        out = np.zeros(PT_CACHE_PARAMS, np.float32)
        inp = None if new is None else np.ascontiguousarray(new, np.float32)
        self._check(self.lib.pt_cache_params(self.h, ptr(out), ptr(inp)), "pt_cache_params")
        return out

    def cache_infer(self, feats_bf16):
        f = np.ascontiguousarray(feats_bf16, np.uint16).reshape(-1, PT_CACHE_INPUTS)
        out = np.zeros((len(f), PT_CACHE_OUTPUTS), np.float32)
        self._check(self.lib.pt_cache_infer(self.h, ptr(f), len(f), ptr(out)), "pt_cache_infer")
        return out

    def cache_train(self, feats_bf16, targets, lr=1e-3):
        f = np.ascontiguousarray(feats_bf16, np.uint16).reshape(-1, PT_CACHE_INPUTS)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, PT_CACHE_OUTPUTS)
        loss = C.c_float()
        grad = np.zeros(PT_CACHE_PARAMS, np.float32)
        self._check(self.lib.pt_cache_train(self.h, ptr(f), ptr(t), len(f), lr, C.byref(loss), ptr(grad)),
                    "pt_cache_train")
        return loss.value, grad

    def cache_configure(self, query_depth=2, train_fraction=0.03, batch=16384, steps=4, lr=1e-3):
        self._check(self.lib.pt_cache_configure(self.h, query_depth, train_fraction, batch, steps, lr),
                    "pt_cache_configure")

    def cache_judge_configure(self, max_ratio=1.0, rows=16384):
        self._check(self.lib.pt_cache_judge_configure(self.h, max_ratio, rows), "pt_cache_judge_configure")

    def cache_records(self, which, max_records=1 << 20):
        """Last batch's training vertices (which=0) or queries (1):
        (features (n, 96) uint16, values (n, 9) float32, slots (n,), total)."""
        cnt = C.c_uint32()
        self._check(self.lib.pt_cache_read_records(self.h, which, 0, None, None, None, C.byref(cnt)),
                    "pt_cache_read_records")
        n = min(cnt.value, max_records)
        feats = np.zeros((n, PT_CACHE_INPUTS), np.uint16)
        vals = np.zeros((n, 9), np.float32)
        slots = np.zeros(n, np.uint32)
        self._check(self.lib.pt_cache_read_records(self.h, which, n, ptr(feats), ptr(vals), ptr(slots),
                                                   C.byref(cnt)), "pt_cache_read_records")
        return feats, vals, slots, cnt.value

    def read_accel(self):
        nn, npm, nb = C.c_uint64(), C.c_uint64(), C.c_uint32()
        self._check(self.lib.pt_read_accel(self.h, None, C.byref(nn), None, None, C.byref(npm),
                                           None, C.byref(nb)), "pt_read_accel")
        nodes = np.zeros((nn.value, 32), np.uint32)
        order = np.zeros(npm.value, np.uint32)
        morton = np.zeros(npm.value, np.uint32)
        info = np.zeros((nb.value, 6), np.uint32)
        self._check(self.lib.pt_read_accel(self.h, ptr(nodes), C.byref(nn), ptr(order), ptr(morton),
                                           C.byref(npm), ptr(info), C.byref(nb)), "pt_read_accel")
        return nodes, order, morton, info


def render_batch(contexts, params, frames, stats=True):
    """pt_render_batch: frames [0, frames[i]) of every contexts[i] in one
    wavefront (contexts[0] runs it); returns the batch's RenderStats."""
    n = len(contexts)
    lib = contexts[0].lib
    hs = (C.c_void_p * n)(*[c.h for c in contexts])
    ps = (RenderParams * n)(*params)
    fs = (C.c_uint32 * n)(*frames)
    s = RenderStats() if stats else None
    contexts[0]._check(lib.pt_render_batch(hs, ps, fs, n, C.byref(s) if s is not None else None),
                       "pt_render_batch")
    return s


def autotune_batch(contexts, params, frames):
    """pt_autotune_batch: the fastest launch shape of this batch (kept in
    contexts[0] for renders whose params[0] has PT_FLAG_TUNED)."""
    n = len(contexts)
    lib = contexts[0].lib
    hs = (C.c_void_p * n)(*[c.h for c in contexts])
    ps = (RenderParams * n)(*params)
    fs = (C.c_uint32 * n)(*frames)
    r = TuneResult()
    contexts[0]._check(lib.pt_autotune_batch(hs, ps, fs, n, C.byref(r)), "pt_autotune_batch")
    return r


def header_symbols():
    """Function names declared in include/*.h."""
    import re
    names = []
    for h in sorted((REPO / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(pt_\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))
