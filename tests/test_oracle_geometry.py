"""Pins the oracle's intersection programs, hit records, instance transforms
and LBVH against closed forms, invariants and brute force.

Reference behaviour followed: sphere.cu:32-93, aarect.cu:20-136,
box.cu:24-88, triangle.cu (replaced by the watertight test of Woop, Benthin
and Wald, JCGT 2(1) 2013), transforms.hpp:51-419, the LBVH of Karras (HPG
2012) collapsed to 4-wide nodes.
"""
import math

import numpy as np
import pytest

import ptamd
import ptref
from helpers import PyScene, icosphere, random_rays

F32 = np.float32


def one_ray(o, d, tmin=1e-4):
    r = np.zeros(1, ptamd.RAY_DTYPE)
    r["org"], r["dir"], r["tmin"], r["tmax"] = o, d, tmin, np.inf
    return r


def trace1(orc, o, d, tmin=1e-4):
    h = orc.trace(one_ray(o, d, tmin))[0]
    return float(h["t"]), int(h["inst"]), int(h["prim"])


# ------------------------------------------------------------ spheres
@pytest.fixture(scope="module")
def unit_sphere():
    s = PyScene()
    m = s.lambert((0.5, 0.5, 0.5))
    s.spheres.append(((0, 0, 0), 1.0, m))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    orc = ptref.OracleScene(s.desc())
    yield orc
    orc.close()


def test_sphere_closed_forms(unit_sphere):
    orc = unit_sphere
    assert trace1(orc, (0, 0, -5), (0, 0, 1)) == (4.0, 0, 0)
    # unnormalised direction: t scales by 1/|d|
    assert trace1(orc, (0, 0, -5), (0, 0, 2))[0] == 2.0
    # origin inside: the second root
    assert trace1(orc, (0, 0, 0), (1, 0, 0))[0] == 1.0
    # ray passing beside the sphere misses; just inside the silhouette hits
    assert trace1(orc, (1.001, 0, -5), (0, 0, 1))[1] == -1
    t, i, _ = trace1(orc, (0.999, 0, -5), (0, 0, 1))
    assert i == 0 and abs(t - (5 - math.sqrt(1 - 0.999 ** 2))) < 1e-5
    # tmin excludes the first root, the second is taken
    t, _, _ = trace1(orc, (0, 0, -5), (0, 0, 1), tmin=4.5)
    assert t == 6.0


def test_sphere_hit_record(unit_sphere):
    orc = unit_sphere
    # (point hit, expected normal, expected u, v); u = 1 - (atan2(z,x)+pi)/2pi,
    # v = (asin(y) + pi/2)/pi  (sphere.cu:65-93)
    cases = [((0, 0, -5), (0, 0, 1), (0, 0, -1), 0.75, 0.5),
             ((5, 0, 0), (-1, 0, 0), (1, 0, 0), 0.5, 0.5),
             ((0, 5, 0), (0, -1, 0), (0, 1, 0), None, 1.0),
             ((0, -5, 0), (0, 1, 0), (0, -1, 0), None, 0.0)]
    for o, d, n, u, v in cases:
        t, inst, prim = trace1(orc, o, d)
        rec, mat = orc.hit_record(inst, prim, t, o, d)
        np.testing.assert_allclose(rec[0:3], n, atol=1e-6)   # P on the unit sphere
        np.testing.assert_allclose(rec[3:6], n, atol=1e-6)   # N = (P - c) / r
        np.testing.assert_allclose(rec[6:9], -np.array(d, F32) / np.linalg.norm(d), atol=1e-7)  # Wo
        if u is not None:
            assert abs(rec[9] - u) < 1e-6
        assert abs(rec[10] - v) < 1e-6
        assert mat == 0


# -------------------------------------------------------------- rects
def test_rect_closed_forms():
    s = PyScene()
    m = s.lambert((1, 1, 1))
    # axis Y rect x in [0,1], z in [0,2] at y = 0.5, flipped normal
    s.rects.append((ptamd.PT_AXIS_Y, 1, 0.0, 1.0, 0.0, 2.0, 0.5, m))
    s.instance(ptamd.PT_GEOM_RECTS, 0, 1)
    orc = ptref.OracleScene(s.desc())
    t, i, p = trace1(orc, (0.25, 2, 0.5), (0, -1, 0))
    assert (t, i, p) == (1.5, 0, 0)
    rec, _ = orc.hit_record(i, p, t, (0.25, 2, 0.5), (0, -1, 0))
    np.testing.assert_array_equal(rec[3:6], F32([0, -1, 0]))  # flip
    assert abs(rec[9] - 0.25) < 1e-7 and abs(rec[10] - 0.25) < 1e-7
    # bounds are inclusive (aarect.cu:20-69)
    assert trace1(orc, (1.0, 2, 2.0), (0, -1, 0))[1] == 0
    assert trace1(orc, (0.0, 2, 0.0), (0, -1, 0))[1] == 0
    assert trace1(orc, (1.0000001, 2, 1), (0, -1, 0))[1] == -1
    assert trace1(orc, (0.5, 2, 2.0000002), (0, -1, 0))[1] == -1
    # parallel ray misses
    assert trace1(orc, (0.5, 0.5, -1), (0, 0, 1))[1] == -1
    orc.close()


# -------------------------------------------------------------- boxes
def test_box_closed_forms():
    s = PyScene()
    m = s.lambert((1, 1, 1))
    s.boxes.append(((0, 0, 0), (1, 2, 3), m))
    s.instance(ptamd.PT_GEOM_BOXES, 0, 1)
    orc = ptref.OracleScene(s.desc())
    # entering through each face: t and the normal (t==t1) - (t==t0)
    faces = [((-1, 0.5, 0.5), (1, 0, 0), 1.0, (-1, 0, 0)), ((2, 0.5, 0.5), (-1, 0, 0), 1.0, (1, 0, 0)),
             ((0.5, -1, 0.5), (0, 1, 0), 1.0, (0, -1, 0)), ((0.5, 4, 0.5), (0, -1, 0), 2.0, (0, 1, 0)),
             ((0.5, 0.5, -2), (0, 0, 1), 2.0, (0, 0, -1)), ((0.5, 0.5, 5), (0, 0, -1), 2.0, (0, 0, 1))]
    for o, d, tt, n in faces:
        t, i, p = trace1(orc, o, d)
        assert (t, i, p) == (tt, 0, 0), (o, d)
        rec, _ = orc.hit_record(i, p, t, o, d)
        np.testing.assert_array_equal(rec[3:6], F32(n))
    # from inside the far slab is offered (box.cu:36-53)
    t, i, p = trace1(orc, (0.5, 0.5, 0.5), (0, 1, 0))
    assert t == 1.5
    rec, _ = orc.hit_record(i, p, t, (0.5, 0.5, 0.5), (0, 1, 0))
    np.testing.assert_array_equal(rec[3:6], F32([0, 1, 0]))
    # a ray through the corner (1, 2, 3) along the diagonal hits at the corner
    d = np.array([1, 2, 3], np.float64)
    o = -d
    t, i, _ = trace1(orc, o, d)
    assert i == 0 and t == 1.0
    orc.close()


# ------------------------------------------------- instance transforms
def rot_y(deg):
    a = math.radians(deg)
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


def test_transformed_box_and_sphere():
    """A box rotated about Y and translated (the Cornell box instances,
    transforms.hpp), hit from several directions: t and P against the same
    geometry evaluated in double precision."""
    s = PyScene()
    m = s.lambert((1, 1, 1))
    s.boxes.append(((0, 0, 0), (165, 330, 165), m))
    R = rot_y(15.0)
    o2w = np.hstack([R, np.array([[265.0], [0.0], [295.0]])])
    s.instance(ptamd.PT_GEOM_BOXES, 0, 1, o2w)
    s.spheres.append(((0, 0, 0), 1.0, m))
    o2w_s = np.hstack([2.0 * np.eye(3), np.array([[0.0], [500.0], [0.0]])])  # uniform scale 2
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1, o2w_s)
    orc = ptref.OracleScene(s.desc())
    centre = R @ np.array([82.5, 165, 82.5]) + np.array([265, 0, 295])
    rng = np.random.default_rng(5)
    for _ in range(50):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o = centre - 1000 * d
        t, i, p = trace1(orc, o, d)
        assert i == 0 and p == 0
        # exact answer: slab test in object space in double precision
        ro = R.T @ (o - np.array([265, 0, 295]))
        rd = R.T @ d
        with np.errstate(divide="ignore"):
            t0 = (np.array([0, 0, 0]) - ro) / rd
            t1 = (np.array([165, 330, 165]) - ro) / rd
        tn = np.max(np.minimum(t0, t1))
        assert abs(t - tn) <= 2e-5 * tn
        rec, _ = orc.hit_record(i, p, t, o, d)
        np.testing.assert_allclose(rec[0:3], o + tn * d, atol=0.02)
        assert abs(np.linalg.norm(rec[3:6]) - 1) < 1e-6
    # scaled sphere: radius 2 around (0, 500, 0)
    t, i, p = trace1(orc, (0, 500, -10), (0, 0, 1))
    assert i == 1 and abs(t - 8.0) < 1e-5
    rec, _ = orc.hit_record(i, p, t, (0, 500, -10), (0, 0, 1))
    np.testing.assert_allclose(rec[3:6], [0, 0, -1], atol=1e-6)
    orc.close()


# ----------------------------------------------------------- triangles
def test_watertight_icosphere():
    """Rays from the centre of a closed mesh toward every vertex and every
    edge midpoint must hit it (no cracks on shared edges and vertices), and
    rays from outside through vertices must hit too."""
    V, Fc = icosphere(3)
    s = PyScene()
    m = s.lambert((1, 1, 1))
    s.meshes.append((V, Fc, None, None, m))
    s.instance(ptamd.PT_GEOM_MESH, 0, 1)
    orc = ptref.OracleScene(s.desc())
    edges = set()
    for a, b, c in Fc:
        for x, y in ((a, b), (b, c), (c, a)):
            edges.add((min(x, y), max(x, y)))
    targets = [V[i].astype(np.float64) for i in range(len(V))]
    targets += [(V[a].astype(np.float64) + V[b]) / 2 for a, b in edges]
    targets = np.array(targets)
    rays = np.zeros(len(targets), ptamd.RAY_DTYPE)
    rays["org"] = 0
    rays["dir"] = targets
    rays["tmin"] = 1e-6
    rays["tmax"] = np.inf
    hits = orc.trace(rays)
    assert np.all(hits["inst"] == 0)
    np.testing.assert_allclose(hits["t"], 1.0, atol=0.01)
    # from outside, aimed through each vertex at the centre
    out = np.zeros(len(V), ptamd.RAY_DTYPE)
    out["org"] = V * 3
    out["dir"] = -V
    out["tmin"] = 1e-6
    out["tmax"] = np.inf
    h2 = orc.trace(out)
    assert np.all(h2["inst"] == 0)
    np.testing.assert_allclose(h2["t"], 2.0, atol=1e-5)
    orc.close()


def test_triangle_hit_record_barycentrics():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    Fc = np.array([[0, 1, 2]], np.uint32)
    N = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1]], np.float32)
    UV = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    s = PyScene()
    m = s.lambert((1, 1, 1))
    s.meshes.append((V, Fc, N, UV, m))
    s.instance(ptamd.PT_GEOM_MESH, 0, 1)
    orc = ptref.OracleScene(s.desc())
    o, d = (0.25, 0.5, 2.0), (0, 0, -1)
    t, i, p = trace1(orc, o, d)
    assert (t, i, p) == (2.0, 0, 0)
    rec, _ = orc.hit_record(i, p, t, o, d)
    np.testing.assert_allclose(rec[0:3], [0.25, 0.5, 0], atol=1e-7)
    np.testing.assert_allclose(rec[3:6], [0, 0, 1], atol=1e-7)
    np.testing.assert_allclose(rec[9:11], [0.25, 0.5], atol=1e-7)
    # a ray through the edge v1-v2 and through a vertex still hits
    assert trace1(orc, (0.5, 0.5, 2), d)[1] == 0
    assert trace1(orc, (1.0, 0.0, 2), d)[1] == 0
    assert trace1(orc, (0.5000001, 0.5000001, 2), d)[1] == -1
    orc.close()


# ----------------------------------------------------------------- LBVH
def spread_bits(v):
    """Bit-by-bit definition of the 10-bit spread (bit i -> bit 3i)."""
    out = np.zeros_like(v, dtype=np.uint32)
    for i in range(10):
        out |= ((v >> np.uint32(i)) & np.uint32(1)) << np.uint32(3 * i)
    return out


def soup_scene(n, seed, r=(0.05, 0.6)):
    rng = np.random.default_rng(seed)
    s = PyScene()
    m = s.lambert((0.5, 0.5, 0.5))
    c = rng.uniform(-10, 10, (n, 3)).astype(np.float32)
    rad = rng.uniform(*r, n).astype(np.float32)
    for i in range(n):
        s.spheres.append((tuple(c[i]), float(rad[i]), m))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, n)
    return s, c, rad


def walk_bvh(nodes, base, count):
    """Leaf ranges and containment of a 4-wide BVH (128-byte nodes: lo.x[4],
    hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4], child[4], count[4])."""
    f = nodes.view(np.float32)
    leaves = []
    stack = [(0, None)]
    seen = 0
    while stack:
        ni, box = stack.pop()
        seen += 1
        assert ni < count
        w = nodes[base + ni]
        fw = f[base + ni]
        for c in range(4):
            child, cnt = int(w[24 + c]), int(w[28 + c])
            if child == 0xFFFFFFFF:
                continue
            lo = np.array([fw[c], fw[8 + c], fw[16 + c]])
            hi = np.array([fw[4 + c], fw[12 + c], fw[20 + c]])
            assert np.all(lo <= hi)
            if box is not None:  # the parent's slot box contains every child slot box
                assert np.all(box[0] <= lo) and np.all(hi <= box[1])
            if cnt > 0:
                leaves.append((child, cnt, lo, hi))
            else:
                stack.append((child, (lo, hi)))
    assert seen == count  # every node reachable exactly once
    return leaves


@pytest.mark.parametrize("n,seed", [(1, 0), (5, 1), (300, 2), (3000, 3)])
def test_bvh_invariants_and_morton(n, seed):
    s, c, rad = soup_scene(n, seed)
    orc = ptref.OracleScene(s.desc())
    nodes, order, morton, info = orc.accel()
    nb, ncnt, pb, pcnt, kind, first = (int(x) for x in info[0])
    assert pcnt == n and kind == ptamd.PT_GEOM_SPHERES
    # order is a permutation and Morton codes are sorted
    assert np.array_equal(np.sort(order[pb:pb + n]), np.arange(n))
    mc = morton[pb:pb + n]
    assert np.all(mc[1:] >= mc[:-1])
    # Morton codes: the magic-number spread equals the bit-by-bit interleave
    # of the 10-bit quantised centroid (sphere bounds c +- r, sphere.cu:96-100)
    lo = c - rad[:, None]
    hi = c + rad[:, None]
    cen = (lo + hi) * F32(0.5)
    clo, chi = cen.min(axis=0), cen.max(axis=0)
    ext = chi - clo
    scale = np.where(ext > 0, F32(1024.0) / np.where(ext > 0, ext, 1), 0).astype(np.float32)
    f = (cen - clo) * scale
    q = np.where(f >= 1023, 1023, np.floor(f)).astype(np.uint32)
    code = (spread_bits(q[:, 0]) << np.uint32(2)) | (spread_bits(q[:, 1]) << np.uint32(1)) | spread_bits(q[:, 2])
    np.testing.assert_array_equal(mc, code[order[pb:pb + n]])
    # stable: equal codes keep primitive order
    eq = mc[1:] == mc[:-1]
    assert np.all(order[pb:pb + n][1:][eq] > order[pb:pb + n][:-1][eq])
    # every primitive in exactly one leaf, leaf boxes contain their bounds
    leaves = walk_bvh(nodes, nb, ncnt)
    covered = np.zeros(n, np.int32)
    for first_i, cnt, llo, lhi in leaves:
        assert cnt <= 4
        covered[first_i:first_i + cnt] += 1
        prims = order[pb + first_i: pb + first_i + cnt]
        assert np.all(llo <= lo[prims].min(axis=0)) and np.all(hi[prims].max(axis=0) <= lhi)
    assert np.all(covered == 1)
    orc.close()


def mixed_scene():
    s = PyScene()
    m = s.lambert((0.5, 0.5, 0.5))
    rng = np.random.default_rng(11)
    for i in range(200):
        s.spheres.append((tuple(rng.uniform(-20, 20, 3)), float(rng.uniform(0.2, 2)), m))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 200)
    for i in range(30):
        axis = int(rng.integers(0, 3))
        a0, b0 = rng.uniform(-20, 10, 2)
        s.rects.append((axis, int(rng.integers(0, 2)), a0, a0 + rng.uniform(1, 10), b0, b0 + rng.uniform(1, 10),
                        float(rng.uniform(-20, 20)), m))
    s.instance(ptamd.PT_GEOM_RECTS, 0, 30)
    for i in range(20):
        lo = rng.uniform(-20, 15, 3)
        s.boxes.append((tuple(lo), tuple(lo + rng.uniform(0.5, 5, 3)), m))
    s.instance(ptamd.PT_GEOM_BOXES, 0, 10)
    for k in range(3):  # the same 10 boxes instanced under rotations
        o2w = np.hstack([rot_y(30.0 * k + 10), rng.uniform(-5, 5, (3, 1))])
        s.instance(ptamd.PT_GEOM_BOXES, 10, 10, o2w)
    V, Fc = icosphere(2, radius=3.0)
    s.meshes.append((V, Fc, None, None, m))
    for k in range(3):
        o2w = np.hstack([rot_y(45.0 * k) * (0.5 + 0.5 * k), rng.uniform(-15, 15, (3, 1))])
        s.instance(ptamd.PT_GEOM_MESH, 0, 1, o2w)
    return s


@pytest.mark.parametrize("name", ["soup", "mixed", "coincident"])
def test_bvh_closest_equals_brute_force(name):
    if name == "soup":
        s, _, _ = soup_scene(2000, 4)
    elif name == "mixed":
        s = mixed_scene()
    else:  # many coincident spheres: ties broken by (instance, primitive)
        s = PyScene()
        m = s.lambert((1, 1, 1))
        for i in range(64):
            s.spheres.append(((0, 0, 0) if i % 2 else (1, 1, 1), 1.0, m))
        s.instance(ptamd.PT_GEOM_SPHERES, 0, 64)
        s.instance(ptamd.PT_GEOM_SPHERES, 0, 64)
    orc = ptref.OracleScene(s.desc())
    ext = 12 if name == "soup" else (25 if name == "mixed" else 3)
    rays = random_rays(20000, -ext, ext, seed=3)
    hb = orc.trace(rays)
    hf = orc.trace(rays, brute=True)
    np.testing.assert_array_equal(hb["inst"], hf["inst"])
    np.testing.assert_array_equal(hb["prim"], hf["prim"])
    np.testing.assert_array_equal(hb["t"], hf["t"])
    assert (hb["inst"] >= 0).mean() > 0.1
    if name == "coincident":
        hit = hb["inst"] >= 0
        assert np.all(hb["inst"][hit] == 0)
        assert set(np.unique(hb["prim"][hit])) <= {0, 1}
    orc.close()
