"""Pins the oracle's open cylinder and moving sphere.

Reference code followed: cylinder.cu:24-107 (intersection fixed to its
intent: the reference's b and c coefficients read d.x P.x + d.z + P.z and
P.x^2 + P.z^2 + R^2, and it rejects t < delta), moving_sphere.cu:57-147
(center lerped over [time0, time1], both roots offered, bounds at time0 and
time1), raygen.cu:69 (the ray time drawn in the camera's shutter interval).

Cylinder: distances along axis-perpendicular rays are (distance to the axis
- R) from outside and R from the axis, the height range is inclusive,
rays parallel to the axis miss, normals are the unit radial direction.
Moving sphere: with center0 == center1 it renders bit for bit as a static
sphere; swept across a pencil of rays over the shutter it is hit by the
fraction of time its center is within one radius of them.
"""
import math

import numpy as np
import pytest

import ptamd
import ptref
from helpers import PyScene, random_rays


def rays(orgs, dirs, tmin=1e-4):
    orgs, dirs = np.atleast_2d(orgs), np.atleast_2d(dirs)
    n = max(len(orgs), len(dirs))
    r = np.zeros(n, ptamd.RAY_DTYPE)
    r["org"] = np.float32(np.broadcast_to(orgs, (n, 3)))
    r["dir"] = np.float32(np.broadcast_to(dirs, (n, 3)))
    r["tmin"] = tmin
    r["tmax"] = np.inf
    return r


@pytest.fixture(scope="module")
def cyl():
    s = PyScene()
    s.cylinders.append(((1.0, 2.0, 3.0), 4.0, 1.5, s.lambert((0.5, 0.5, 0.5))))
    s.instance(ptamd.PT_GEOM_CYLINDERS, 0, 1)
    orc = ptref.OracleScene(s.desc())
    yield orc
    orc.close()


def test_cylinder_closed_forms(cyl):
    # from -x at heights inside [2, 6]: entry at x = 1 - 1.5, t = 9.5
    ys = np.float32([2.0, 2.5, 4.0, 5.999, 6.0])
    h = cyl.trace(rays(np.stack([np.full(5, -10.0), ys, np.full(5, 3.0)], 1), (1, 0, 0)))
    np.testing.assert_array_equal(h["inst"], 0)
    np.testing.assert_allclose(h["t"], 9.5, rtol=1e-6)
    # above the top, below the base, parallel to the axis
    h = cyl.trace(rays([[-10, 6.01, 3], [-10, 1.99, 3], [1.5, -10, 3], [1, 20, 3]],
                       [[1, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0]]))
    assert (h["inst"] < 0).all()
    # from the axis: the inner wall at distance R in every horizontal direction
    a = np.linspace(0, 2 * np.pi, 17)[:-1]
    d = np.stack([np.cos(a), np.zeros_like(a), np.sin(a)], 1)
    h = cyl.trace(rays((1, 4, 3), d))
    np.testing.assert_allclose(h["t"], 1.5, rtol=2e-6)
    # a direction with a y component leaves through the top instead
    h = cyl.trace(rays((1, 4, 3), (1, 3, 0)))
    assert h["inst"][0] < 0


def test_cylinder_hits_and_normals(cyl):
    r = random_rays(20000, -6, 10, seed=3)
    h = cyl.trace(r)
    hb = cyl.trace(r, brute=True)
    assert np.array_equal(h["inst"], hb["inst"]) and np.array_equal(h["t"], hb["t"])
    hit = np.flatnonzero(h["inst"] >= 0)
    assert len(hit) > 500
    o = r["org"][hit].astype(np.float64)
    d = r["dir"][hit].astype(np.float64)
    P = o + h["t"][hit, None] * d
    rad = np.hypot(P[:, 0] - 1, P[:, 2] - 3)
    np.testing.assert_allclose(rad, 1.5, rtol=1e-4)
    assert (P[:, 1] >= 2 - 1e-4).all() and (P[:, 1] <= 6 + 1e-4).all()
    for k in hit[:200]:
        rec, _ = cyl.hit_record(0, 0, h["t"][k], r["org"][k], r["dir"][k])
        Pk, N = rec[:3].astype(np.float64), rec[3:6].astype(np.float64)
        want = np.array([Pk[0] - 1, 0, Pk[2] - 3]) / np.hypot(Pk[0] - 1, Pk[2] - 3)
        np.testing.assert_allclose(N, want, atol=2e-6)


def moving_scene(c0, c1, t0=0.0, t1=1.0, r=0.5, moving=True, material=None):
    s = PyScene()
    m = material if material is not None else s.lambert((0.7, 0.6, 0.5))
    if moving:
        s.moving.append((c0, c1, r, t0, t1, m))
        s.instance(ptamd.PT_GEOM_MOVING_SPHERES, 0, 1)
    else:
        s.spheres.append((c0, r, m))
        s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    s.spheres.append(((0, -100.5, 0), 100.0, s.lambert((0.5, 0.5, 0.5))))
    s.instance(ptamd.PT_GEOM_SPHERES, len(s.spheres) - 1, 1)
    return s


def test_moving_sphere_queries_at_time_zero():
    # time range [-1, 1]: at time 0 the center is the midpoint (2, 0, 0)
    orc = ptref.OracleScene(moving_scene((0, 0, 0), (4, 0, 0), -1.0, 1.0).desc())
    h = orc.trace(rays([[2, 0, -10], [0, 0, -10]], (0, 0, 1)))
    assert h["inst"][0] == 0 and h["t"][0] == pytest.approx(9.5, rel=1e-6)
    assert h["inst"][1] < 0 or h["inst"][1] == 1
    r = random_rays(20000, -3, 6, seed=4)
    h, hb = orc.trace(r), orc.trace(r, brute=True)
    assert np.array_equal(h["inst"], hb["inst"]) and np.array_equal(h["t"], hb["t"])
    orc.close()


def look(origin, target, W, H, half=0.001):
    cam = ptamd.Camera()
    o, t = np.array(origin, np.float64), np.array(target, np.float64)
    w = (o - t) / np.linalg.norm(o - t)
    u = np.cross((0, 1, 0), w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    cam.lower_left[:] = list(t - half * u - half * v)
    cam.horizontal[:] = list(2 * half * u)
    cam.vertical[:] = list(2 * half * v)
    cam.origin[:] = list(o)
    cam.u[:], cam.v[:], cam.w[:] = list(u), list(v), list(w)
    cam.lens_radius, cam.time0, cam.time1 = 0.0, 0.0, 1.0
    return cam


def test_static_moving_sphere_renders_as_a_sphere():
    W, H = 32, 32
    cam = look((0, 1, 6), (0, 0, 0), W, H, half=2.0)
    params = ptamd.RenderParams(W, H, 8, 1, 3, 1, ptamd.PT_FLAG_KEEP_PATHS, 0)
    out = []
    for moving in (True, False):
        orc = ptref.OracleScene(moving_scene((0.2, 0.1, 0), (0.2, 0.1, 0), moving=moving).desc())
        acc, (rgb, sig, b), _ = orc.render(cam, params, 0, 2, keep_last=True)
        out.append((acc, rgb, sig))
        orc.close()
    # accumulation and per-path radiance (the signatures differ: the ground
    # sphere's primitive index is 0 in one scene and 1 in the other)
    for a, b in zip(out[0][:2], out[1][:2]):
        assert np.array_equal(a, b)


def test_swept_sphere_is_hit_for_its_share_of_the_shutter():
    """center x = -2 + 4 t over t in [0, 1]; a pencil of rays along -z
    through the origin meets it while |x| < 0.5: a quarter of the shutter"""
    W, H = 64, 64
    s = PyScene()
    m = s.material(ptamd.PT_MAT_NORMAL, -1, params=(1.0,))
    s.moving.append(((-2, 0, 0), (2, 0, 0), 0.5, 0.0, 1.0, m))
    s.instance(ptamd.PT_GEOM_MOVING_SPHERES, 0, 1)
    s.miss = (ptamd.PT_MISS_CONSTANT, (0, 0, 0), (0, 0, 0), -1, 1)
    orc = ptref.OracleScene(s.desc())
    cam = look((0, 0, 50), (0, 0, 0), W, H)
    params = ptamd.RenderParams(W, H, 4, 1, 3, 1, ptamd.PT_FLAG_KEEP_PATHS, 0)
    hits = 0
    n = 0
    for f in range(3):
        _, (rgb, _, _), _ = orc.render(cam, params, f, 1, keep_last=True)
        hits += int((rgb.sum(1) > 0).sum())
        n += W * H
    p = 0.25
    assert abs(hits - n * p) <= 4.5 * math.sqrt(n * p * (1 - p))
    orc.close()
