"""Test helpers: scene descriptions built directly in Python (ctypes views of
include/pt_api.h), ray sets and comparison utilities."""
from __future__ import annotations

import ctypes as C

import numpy as np

import ptamd


class PyScene:
    """Owns the arrays behind a ptamd.SceneDesc."""

    def __init__(self):
        self.spheres, self.rects, self.boxes = [], [], []
        self.volumes = []      # (shape, a (3), b (3), density, mat)
        self.cylinders = []    # (origin (3), length, radius, mat)
        self.moving = []       # (center0 (3), center1 (3), radius, time0, time1, mat)
        self.meshes = []       # (verts (n,3) f32, idx (m,3) u32, normals|None, uvs|None, mat)
        self.instances = []    # (kind, first, count, o2w 3x4 or None)
        self.materials = []    # (type, tex0, tex1, params)
        self.textures = []     # (type, rgb, image (h,w,4) or None)
        self.lights = []       # dict
        self.miss = (ptamd.PT_MISS_GRADIENT, (1, 1, 1), (0.5, 0.7, 1.0), -1, 1)
        self._keep = []

    def texture(self, rgb=None, image=None):
        self.textures.append((ptamd.PT_TEX_IMAGE if image is not None else ptamd.PT_TEX_CONSTANT,
                              rgb or (0, 0, 0), image, (0, 0), None))
        return len(self.textures) - 1

    def checker(self, odd, even):
        """odd where sin(10x) sin(10y) sin(10z) < 0, else even (texture indices)"""
        self.textures.append((ptamd.PT_TEX_CHECKER, (0, 0, 0), None, (odd, even), None))
        return len(self.textures) - 1

    def noise(self, scale, ranvec, perms):
        """ranvec (256, 3) gradient vectors, perms (3, 256) permutations"""
        data = np.zeros(ptamd.PT_NOISE_FLOATS, np.float32)
        data[:1024].reshape(256, 4)[:, :3] = ranvec
        data[1024:] = np.asarray(perms, np.float32).ravel()
        self.textures.append((ptamd.PT_TEX_NOISE, (scale, 0, 0), None, (0, 0), data))
        return len(self.textures) - 1

    def gradient(self, a, b, c):
        data = np.zeros(12, np.float32)
        data.reshape(3, 4)[:, :3] = [a, b, c]
        self.textures.append((ptamd.PT_TEX_GRADIENT, (0, 0, 0), None, (0, 0), data))
        return len(self.textures) - 1

    def material(self, mtype, tex0, tex1=-1, params=(0, 0, 0, 0)):
        self.materials.append((mtype, tex0, tex1, tuple(params) + (0,) * (4 - len(params))))
        return len(self.materials) - 1

    def lambert(self, rgb):
        return self.material(ptamd.PT_MAT_LAMBERTIAN, self.texture(rgb))

    def instance(self, kind, first, count, o2w=None):
        self.instances.append((kind, first, count, o2w))

    def desc(self) -> ptamd.SceneDesc:
        d = ptamd.SceneDesc()
        keep = self._keep
        keep.clear()

        def arr(T, items):
            a = (T * max(1, len(items)))()
            for i, it in enumerate(items):
                a[i] = it
            keep.append(a)
            return a

        d.num_spheres = len(self.spheres)
        d.spheres = arr(ptamd.SphereDesc, [ptamd.SphereDesc((C.c_float * 3)(*c), r, m)
                                           for c, r, m in self.spheres])
        d.num_rects = len(self.rects)
        d.rects = arr(ptamd.RectDesc, [ptamd.RectDesc(*r) for r in self.rects])
        d.num_volumes = len(self.volumes)
        d.volumes = arr(ptamd.VolumeDesc, [ptamd.VolumeDesc(sh, (C.c_float * 3)(*a), (C.c_float * 3)(*b), den, m)
                                           for sh, a, b, den, m in self.volumes])
        d.num_cylinders = len(self.cylinders)
        d.cylinders = arr(ptamd.CylinderDesc, [ptamd.CylinderDesc((C.c_float * 3)(*o), L, R, m)
                                               for o, L, R, m in self.cylinders])
        d.num_moving_spheres = len(self.moving)
        d.moving_spheres = arr(ptamd.MovingSphereDesc,
                               [ptamd.MovingSphereDesc((C.c_float * 3)(*c0), (C.c_float * 3)(*c1), r, t0, t1, m)
                                for c0, c1, r, t0, t1, m in self.moving])
        d.num_boxes = len(self.boxes)
        d.boxes = arr(ptamd.BoxDesc, [ptamd.BoxDesc((C.c_float * 3)(*lo), (C.c_float * 3)(*hi), m)
                                      for lo, hi, m in self.boxes])
        ms = []
        for mesh in self.meshes:
            v, idx, n, uv, mat = mesh[:5]
            fi = mesh[5] if len(mesh) > 5 else None  # per-face vector-texture indices
            v = np.ascontiguousarray(v, np.float32)
            idx = np.ascontiguousarray(idx, np.uint32)
            keep += [v, idx]
            md = ptamd.MeshDesc()
            md.num_vertices = len(v)
            md.vertices = v.ctypes.data_as(C.POINTER(C.c_float))
            if n is not None:
                n = np.ascontiguousarray(n, np.float32)
                keep.append(n)
                md.normals = n.ctypes.data_as(C.POINTER(C.c_float))
            if uv is not None:
                uv = np.ascontiguousarray(uv, np.float32)
                keep.append(uv)
                md.uvs = uv.ctypes.data_as(C.POINTER(C.c_float))
            md.num_triangles = len(idx)
            md.indices = idx.ctypes.data_as(C.POINTER(C.c_uint32))
            md.material = mat
            if fi is not None:
                fi = np.ascontiguousarray(fi, np.int32)
                keep.append(fi)
                md.face_indices = fi.ctypes.data_as(C.POINTER(C.c_int32))
            ms.append(md)
        d.num_meshes = len(ms)
        d.meshes = arr(ptamd.MeshDesc, ms)
        ins = []
        for kind, first, count, o2w in self.instances:
            x = ptamd.InstanceDesc()
            x.kind, x.first, x.count = kind, first, count
            if o2w is None:
                m = np.eye(4)[:3]
                x.has_transform = 0
            else:
                m = np.asarray(o2w, np.float64).reshape(3, 4)
                x.has_transform = 1
            full = np.vstack([m, [0, 0, 0, 1]])
            inv = np.linalg.inv(full)[:3]
            x.obj_to_world = (C.c_float * 12)(*m.astype(np.float32).ravel())
            x.world_to_obj = (C.c_float * 12)(*inv.astype(np.float32).ravel())
            ins.append(x)
        d.num_instances = len(ins)
        d.instances = arr(ptamd.InstanceDesc, ins)
        mats = []
        for t, t0, t1, p in self.materials:
            m = ptamd.MaterialDesc()
            m.type = t
            m.tex = (C.c_int32 * 2)(t0, t1)
            m.param = (C.c_float * 4)(*p)
            mats.append(m)
        d.num_materials = len(mats)
        d.materials = arr(ptamd.MaterialDesc, mats)
        texs = []
        for t, rgb, img, wh, data in self.textures:
            x = ptamd.TextureDesc()
            x.type = t
            x.rgb = (C.c_float * 3)(*rgb)
            x.width, x.height = wh
            if img is not None:
                data = img
                x.height, x.width = img.shape[:2]
            if data is not None:
                data = np.ascontiguousarray(data, np.float32)
                keep.append(data)
                x.texels = data.ctypes.data_as(C.POINTER(C.c_float))
            texs.append(x)
        d.num_textures = len(texs)
        d.textures = arr(ptamd.TextureDesc, texs)
        lts = []
        for l in self.lights:
            x = ptamd.LightDesc()
            x.type = l["type"]
            x.axis = l.get("axis", 0)
            x.a0, x.a1, x.b0, x.b1, x.k = l.get("rect", (0, 0, 0, 0, 0))
            x.center = (C.c_float * 3)(*l.get("center", (0, 0, 0)))
            x.radius = l.get("radius", 0)
            x.emission = (C.c_float * 3)(*l["emission"])
            lts.append(x)
        d.num_lights = len(lts)
        d.lights = arr(ptamd.LightDesc, lts)
        mt, c1, c2, tex, sph = self.miss
        d.miss = ptamd.MissDesc(mt, (C.c_float * 3)(*c1), (C.c_float * 3)(*c2), tex, sph)
        self._desc = d
        return d


def icosphere(subdiv: int, radius=1.0, center=(0, 0, 0)):
    """Unindexed-free icosphere: (verts (n,3), tris (m,3)) with 20*4^subdiv faces."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    cache = {}

    def mid(a, b):
        key = (min(a, b), max(a, b))
        if key not in cache:
            m = verts[a] + verts[b]
            verts.append(m / np.linalg.norm(m))
            cache[key] = len(verts) - 1
        return cache[key]

    for _ in range(subdiv):
        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    V = np.array(verts) * radius + np.array(center)
    return V.astype(np.float32), np.array(f, np.uint32)


def random_rays(n, lo, hi, seed=1, tmin=1e-4):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, ptamd.RAY_DTYPE)
    r["org"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    r["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["tmin"] = tmin
    r["tmax"] = np.inf
    return r


def camera_rays(cam, W, H, jitter=0.5):
    """Pinhole rays through pixel (x + jitter, y + jitter), tmin 1e-6."""
    r = np.zeros(W * H, ptamd.RAY_DTYPE)
    ys, xs = np.divmod(np.arange(W * H), W)
    u = ((xs.astype(np.float32) + np.float32(jitter)) / np.float32(W))[:, None]
    v = ((ys.astype(np.float32) + np.float32(jitter)) / np.float32(H))[:, None]
    llc, hor, ver, org = (np.array(getattr(cam, k)[:], np.float32)
                          for k in ("lower_left", "horizontal", "vertical", "origin"))
    r["org"] = org
    r["dir"] = ((llc + u * hor) + v * ver) - org
    r["tmin"] = 1e-6
    r["tmax"] = np.inf
    return r


def same_hits(a, b):
    return np.array_equal(a["inst"], b["inst"]) and np.array_equal(a["prim"], b["prim"])
