"""Pins the oracle's constant-density media, isotropic phase function and
normal shader, and the Final_Next_Week scene's host construction.

Reference code followed: volume_sphere.cu:21-115 and volume_box.cu:20-94
(boundary entry and exit, free flight -ln(xi)/density, bounds),
isotropic.cu:27-51 and isotropic.cuh (uniform sphere sampling; the
reference's pi/4 pdf and value are replaced by the phase function 1/(4 pi)),
normal_shader.cu:22-34, scenes.hpp:263-375.

Media are pinned by what the physics fixes: a ray crossing a chord of
length L scatters with probability 1 - exp(-density L), its scatter depth
follows the exponential truncated at L (mean 1/s - L e^(-sL) / (1 - e^(-sL))),
a ray starting inside only crosses the rest of the chord, and a ray that
misses the boundary never scatters.  The free-flight xi of a ray query is a
hash of the ray's index, so n rays are n independent trials.
"""
import math

import numpy as np
import pytest

import ptamd
import ptref
from helpers import PyScene


def volume_scene(shape, a, b, density, xform=None):
    s = PyScene()
    m = s.material(ptamd.PT_MAT_ISOTROPIC, s.texture((0.5, 0.5, 0.5)))
    s.volumes.append((shape, a, b, density, m))
    s.instance(ptamd.PT_GEOM_VOLUMES, 0, 1, xform)
    return s


def rays_along(n, org, d, tmin=1e-3):
    r = np.zeros(n, ptamd.RAY_DTYPE)
    r["org"] = np.float32(org)
    r["dir"] = np.float32(d)
    r["tmin"] = tmin
    r["tmax"] = np.inf
    return r


def binom_ok(k, n, p, z=4.5):
    return abs(k - n * p) <= z * math.sqrt(n * p * (1 - p)) + 1


def trunc_exp_mean(s, L):
    e = math.exp(-s * L)
    return 1 / s - L * e / (1 - e)


@pytest.mark.parametrize("shape", [ptamd.PT_VOLUME_SPHERE, ptamd.PT_VOLUME_BOX])
@pytest.mark.parametrize("dlen", [1.0, 2.5])
def test_transmittance_through_chord(shape, dlen):
    R, dens, n = 2.0, 0.3, 40000
    if shape == ptamd.PT_VOLUME_SPHERE:
        sc = volume_scene(shape, (1, 2, 3), (R, 0, 0), dens)
        L = 2 * R
        org, entry = (1, 2, 3 - 10), 10 - R
    else:
        sc = volume_scene(shape, (-1, -1, -1), (3, 3, 3), dens)
        L = 4.0
        org, entry = (1, 1, -10), 9.0
    orc = ptref.OracleScene(sc.desc())
    h = orc.trace(rays_along(n, org, (0, 0, dlen)))
    hit = h["inst"] >= 0
    assert binom_ok(hit.sum(), n, 1 - math.exp(-dens * L))
    depth = h["t"][hit].astype(np.float64) * dlen - entry  # distance past the entry
    assert depth.min() >= -1e-4 and depth.max() <= L + 1e-4
    m, sd = depth.mean(), depth.std() / math.sqrt(hit.sum())
    assert abs(m - trunc_exp_mean(dens, L)) < 5 * sd
    # brute force over the primitives gives the same hits
    hb = orc.trace(rays_along(n, org, (0, 0, dlen)), brute=True)
    assert np.array_equal(h["inst"], hb["inst"]) and np.array_equal(h["t"], hb["t"])
    orc.close()


def test_ray_starting_inside_crosses_the_rest():
    R, dens, n = 3.0, 0.25, 40000
    orc = ptref.OracleScene(volume_scene(ptamd.PT_VOLUME_SPHERE, (0, 0, 0), (R, 0, 0), dens).desc())
    h = orc.trace(rays_along(n, (0, 0, 0), (1, 0, 0)))
    hit = h["inst"] >= 0
    assert binom_ok(hit.sum(), n, 1 - math.exp(-dens * R))
    t = h["t"][hit].astype(np.float64)
    assert t.min() > 1e-3 and t.max() < R
    assert abs(t.mean() - 1e-3 - trunc_exp_mean(dens, R - 1e-3)) < 5 * t.std() / math.sqrt(hit.sum())
    orc.close()


def test_missing_and_transformed_volumes():
    n = 5000
    orc = ptref.OracleScene(volume_scene(ptamd.PT_VOLUME_SPHERE, (0, 0, 0), (1, 0, 0), 50.0).desc())
    assert (orc.trace(rays_along(n, (0, 1.01, -5), (0, 0, 1)))["inst"] < 0).all()  # passes above
    assert (orc.trace(rays_along(n, (0, 0, 5), (0, 0, 1)))["inst"] < 0).all()      # points away
    orc.close()
    # a unit box scaled by 4 along x and moved: the world chord is 4 long,
    # the optical depth is density x the object-space chord (1), as the
    # intersection runs on the object-space ray
    M = np.array([[4, 0, 0, 10], [0, 1, 0, 0], [0, 0, 1, 0]], np.float64)
    dens = 0.2
    orc = ptref.OracleScene(volume_scene(ptamd.PT_VOLUME_BOX, (0, 0, 0), (1, 1, 1), dens, M).desc())
    h = orc.trace(rays_along(40000, (0, 0.5, 0.5), (1, 0, 0)))
    hit = h["inst"] >= 0
    assert binom_ok(hit.sum(), 40000, 1 - math.exp(-dens * 1))
    t = h["t"][hit]
    assert t.min() >= 10 - 1e-4 and t.max() <= 14 + 1e-4
    orc.close()


def test_dense_medium_scatters_near_entry_and_surfaces_win():
    """A surface inside a medium is hit when the free flight reaches past it"""
    s = PyScene()
    iso = s.material(ptamd.PT_MAT_ISOTROPIC, s.texture((1, 1, 1)))
    s.volumes.append((ptamd.PT_VOLUME_SPHERE, (0, 0, 0), (5, 0, 0), 0.5, iso))
    s.instance(ptamd.PT_GEOM_VOLUMES, 0, 1)
    s.spheres.append(((0, 0, 0), 1.0, s.lambert((1, 1, 1))))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    orc = ptref.OracleScene(s.desc())
    n = 40000
    h = orc.trace(rays_along(n, (0, 0, -10), (0, 0, 1)))
    surface = h["inst"] == 1
    # the surface at depth 4 into the medium is reached with probability e^(-2)
    assert binom_ok(surface.sum(), n, math.exp(-0.5 * 4))
    np.testing.assert_allclose(h["t"][surface], 9.0, rtol=1e-6)
    assert (h["t"][~surface] < 9.0).all()
    orc.close()


# --------------------------------------------------------- phase function
def test_isotropic_phase_function_and_sampling():
    rng = np.random.default_rng(3)
    c = (0.2, 0.4, 0.9)
    k = 0.25 / math.pi
    dirs = []
    seed = 12345
    for _ in range(4000):
        wo = rng.normal(size=3)
        wo /= np.linalg.norm(wo)
        n = (1.0, 0.0, 0.0)
        s0 = seed
        wi, f, pdf, seed = ptref.brdf_sample_eval(ptamd.PT_MAT_ISOTROPIC, c, (0, 0, 0), 0, 0, wo, n, seed)
        assert pdf == pytest.approx(k, rel=1e-6)
        np.testing.assert_allclose(f, np.float32(c) * np.float32(k), rtol=1e-6)
        assert abs(np.linalg.norm(wi) - 1) < 1e-5
        # three draws: z, phi and the radius that unit_vector cancels
        x = s0
        for _ in range(3):
            x = (1664525 * x + 1013904223) & 0xFFFFFFFF
        assert seed == x
        dirs.append(wi)
    d = np.array(dirs)
    # uniform on the sphere: mean 0, second moments 1/3, both hemispheres of +X
    assert np.abs(d.mean(0)).max() < 0.05
    np.testing.assert_allclose((d ** 2).mean(0), 1 / 3, atol=0.03)
    assert 0.45 < (d[:, 0] > 0).mean() < 0.55
    # evaluation does not depend on the directions
    f, pdf = ptref.brdf_eval(ptamd.PT_MAT_ISOTROPIC, c, (0, 0, 0), 0, 0, (0, 0, 1), (0, 0, -1), (1, 0, 0))
    assert pdf == pytest.approx(k, rel=1e-6)


# ----------------------------------------------------------- normal shader
def normal_scene(use_shading):
    """A unit square in the z = 0 plane (geometric normal +Z toward the
    camera) whose vertex normals all point along +Y"""
    s = PyScene()
    m = s.material(ptamd.PT_MAT_NORMAL, -1, params=(1.0 if use_shading else 0.0,))
    v = np.float32([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]])
    f = np.uint32([[0, 1, 2], [0, 2, 3]])
    nrm = np.float32([[0, 1, 0]] * 4)
    s.meshes.append((v, f, nrm, None, m))
    s.instance(ptamd.PT_GEOM_MESH, 0, 0)
    s.spheres.append(((3, 0, 0), 0.8, m))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    s.miss = (ptamd.PT_MISS_CONSTANT, (0, 0, 0), (0, 0, 0), -1, 1)
    return s


def front_camera():
    """pinhole at (0.5, 0, 4) looking down -z over x in [-1.5, 4.5]"""
    cam = ptamd.Camera()
    cam.lower_left[:] = [-1.5, -1.5, -1]
    cam.horizontal[:] = [6, 0, 0]
    cam.vertical[:] = [0, 3, 0]
    cam.origin[:] = [0.5, 0, 4]
    cam.u[:] = [1, 0, 0]
    cam.v[:] = [0, 1, 0]
    cam.w[:] = [0, 0, 1]
    return cam


@pytest.mark.parametrize("use_shading", [True, False])
def test_normal_shader_colours(use_shading):
    s = normal_scene(use_shading)
    orc = ptref.OracleScene(s.desc())
    W, H = 48, 24
    cam = front_camera()
    params = ptamd.RenderParams(W, H, 8, 1, 3, 1, ptamd.PT_FLAG_KEEP_PATHS, 0)
    _, (rgb, sig, bounces), _ = orc.render(cam, params, 0, 1, keep_last=True)
    rgb = rgb.astype(np.float64)
    n = rgb * 2 - 1
    bg = np.all(rgb == 0, 1)  # the black background
    lit = ~bg
    assert lit.sum() > 200 and bg.sum() > 50
    np.testing.assert_allclose(np.linalg.norm(n[lit], axis=1), 1, atol=1e-5)
    assert (bounces[lit] == 1).all()
    # the square: +Y shading normal or +Z geometric normal; the sphere: a
    # normal facing the camera
    want = np.array([0, 1, 0]) if use_shading else np.array([0, 0, 1])
    sq = lit & np.all(np.isclose(n, want, atol=1e-5), 1)
    assert sq.sum() > 100
    other = lit & ~sq
    # sphere points P = c + 0.8 N seen from the camera: dot(N, eye - P) > 0
    P = np.array([3.0, 0, 0]) + 0.8 * n[other]
    assert other.sum() > 30 and ((n[other] * (np.array([0.5, 0, 4]) - P)).sum(1) > 0).all()
    orc.close()


# ------------------------------------------------------------ the scene
class HostRnd:
    def __init__(self, seed):
        self.g = np.random.MT19937()
        self.g._legacy_seeding(seed)

    def __call__(self):
        return np.float32(int(self.g.random_raw()) >> 8) * np.float32(2.0 ** -24)


def test_next_week_scene():
    sc = ptamd.Scene(ptamd.PT_SCENE_NEXT_WEEK, 50, 50)
    d = sc.desc
    assert d.num_boxes == 400 and d.num_spheres == 6 + 1000 and d.num_rects == 1 and d.num_volumes == 2
    vols = [d.volumes[i] for i in range(2)]
    assert [v.shape for v in vols] == [ptamd.PT_VOLUME_SPHERE] * 2
    assert sorted(float(v.density) for v in vols) == pytest.approx([1e-4, 0.2])
    for v in vols:
        assert d.materials[v.material].type == ptamd.PT_MAT_ISOTROPIC
    # ground box heights: 100 (rnd + 0.01) in draw order
    r = HostRnd(0)
    heights = [np.float32(100) * (r() + np.float32(0.01)) for _ in range(400)]
    got = sorted(float(d.boxes[i].pmax[1]) for i in range(400))
    np.testing.assert_allclose(got, sorted(float(h) for h in heights), rtol=1e-7)
    # the 1000 spheres share one transformed instance
    xf = [d.instances[i] for i in range(d.num_instances) if d.instances[i].has_transform]
    assert len(xf) == 1 and xf[0].kind == ptamd.PT_GEOM_SPHERES and xf[0].count == 1000
    assert sc.info.max_depth == 50


def test_next_week_oracle_render():
    sc = ptamd.Scene(ptamd.PT_SCENE_NEXT_WEEK, 20, 20)
    orc = ptref.OracleScene(sc.desc)
    acc, _, st = orc.render(sc.camera, sc.render_params(), 0, 1)
    assert np.isfinite(acc).all() and acc.max() > 0
    assert st[2] > 0  # shadow rays traced through the enclosing medium
    orc.close()
