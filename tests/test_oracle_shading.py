"""Pins the oracle's BRDFs, sampling frames, lights and miss programs with
closed forms and invariants.

Reference code followed: lambertian.cuh:7-42, oren_nayar.cuh:22-65,
material.cuh:30-63 (schlick, Refract), microfacets.cuh:85-99 (GGX_D),
light_sample.cuh:27-33 (PowerHeuristic), rect_pdf.cu:11-146,
sphere_pdf.cu:8-74, miss.cu:30-95; the tangent frame of Duff et al. (JCGT
6(1) 2017) and Malley's cosine sampling replace the OptiX SDK helpers.
"""
import math

import numpy as np
import pytest

import ptamd
import ptref
from helpers import PyScene

PI = math.pi


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def rand_dirs(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


# ------------------------------------------------------------ frames
def test_onb_orthonormal_right_handed():
    ns = list(rand_dirs(500, 1)) + [np.float32(v) for v in ([0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0],
                                                              [0, -1, 0], unit([1e-4, 0, -1]))]
    for n in ns:
        n = np.asarray(n, np.float32)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        t, b, nn = ptref.onb(n).astype(np.float64)
        M = np.stack([t, b, nn])
        np.testing.assert_allclose(M @ M.T, np.eye(3), atol=2e-6)
        np.testing.assert_allclose(np.cross(t, b), nn, atol=2e-6)
        np.testing.assert_array_equal(nn, n.astype(np.float64))


def test_cosine_sampling_is_malley():
    """Malley's method: cos(theta) = sqrt(1 - u1), azimuth 2 pi u2 in the
    (t, b) plane of the frame; so P(cos <= c) = c^2 and E[cos] = 2/3."""
    g = (np.arange(24) + 0.5) / 24
    for n in rand_dirs(4, 2):
        t, b, nn = ptref.onb(n).astype(np.float64)
        cos = []
        for u1 in g:
            for u2 in g:
                w = ptref.cosine_sample(float(u1), float(u2), n).astype(np.float64)
                assert abs(np.linalg.norm(w) - 1) < 2e-6
                c = w @ nn
                assert abs(c - math.sqrt(1 - np.float32(u1))) < 2e-6
                r = math.sqrt(np.float32(u1))
                assert abs(w @ t - r * math.cos(2 * PI * np.float32(u2))) < 2e-6
                assert abs(w @ b - r * math.sin(2 * PI * np.float32(u2))) < 2e-6
                cos.append(c)
        assert abs(np.mean(cos) - 2 / 3) < 2e-3


# ------------------------------------------------------------- BRDFs
def test_lambertian_closed_form():
    rng = np.random.default_rng(3)
    rho = np.float32([0.7, 0.5, 0.2])
    for _ in range(300):
        n, wo, wi = rand_dirs(3, int(rng.integers(1 << 30)))
        wi = wi * np.float32(rng.uniform(0.5, 3))  # eval normalises Wi
        f, pdf = ptref.brdf_eval(ptamd.PT_MAT_LAMBERTIAN, rho, rho, 0, 0, wo, wi, n)
        c = unit(wi) @ n
        if c < 0:
            np.testing.assert_array_equal(f, 0)
            assert pdf == 1.0  # early return keeps the incoming pdf, which starts at 1
        else:
            np.testing.assert_allclose(f, rho * c / PI, rtol=2e-6, atol=1e-8)
            assert abs(pdf - c / PI) <= 2e-6 * c / PI + 1e-8


def test_lambertian_sample_eval_consistent():
    rho = np.float32([1, 1, 1])
    n = unit([0.3, 0.9, -0.2]).astype(np.float32)
    wo = unit([0.1, 1, 0.4]).astype(np.float32)
    seed = 99
    for _ in range(200):
        wi, f, pdf, seed = ptref.brdf_sample_eval(ptamd.PT_MAT_LAMBERTIAN, rho, rho, 0, 0, wo, n, seed)
        c = wi.astype(np.float64) @ n
        assert c >= -1e-6
        if pdf > 0:
            # f / pdf == rho to an ulp (the value divides rho cos by pi as a
            # reciprocal and products): the furnace property
            np.testing.assert_allclose(f / np.float32(pdf), rho, rtol=2.5e-7, atol=0)


def test_oren_nayar_sigma_zero_is_rho_over_pi_squared():
    """sigma = 0 gives A = 1/pi, B = 0 (materials.hpp:191-215), and the
    reference's evaluation rho/pi * (A + ...) becomes rho / pi^2."""
    rho = np.float32([0.8, 0.4, 0.1])
    for n, wo, wi in zip(rand_dirs(100, 4), rand_dirs(100, 5), rand_dirs(100, 6)):
        f, _ = ptref.brdf_eval(ptamd.PT_MAT_OREN_NAYAR, rho, rho, np.float32(1 / PI), 0.0, wo, wi, n)
        np.testing.assert_allclose(f, rho / PI ** 2, rtol=3e-7)


def test_schlick_endpoints_and_monotone():
    for ior in (1.0, 1.33, 1.5, 2.4):
        r0 = ((1 - ior) / (1 + ior)) ** 2
        assert abs(ptref.schlick(1.0, ior) - r0) < 1e-7
        assert ptref.schlick(0.0, ior) == 1.0
        vals = [ptref.schlick(c, ior) for c in np.linspace(0, 1, 101)]
        assert all(a >= b for a, b in zip(vals, vals[1:]))


def test_refract_snell_and_tir():
    rng = np.random.default_rng(7)
    for _ in range(2000):
        n = rand_dirs(1, int(rng.integers(1 << 30)))[0]
        v = rand_dirs(1, int(rng.integers(1 << 30)))[0] * np.float32(rng.uniform(0.2, 5))
        if unit(v) @ n > 0:
            v = -v  # incoming against the normal
        eta = float(np.float32(rng.uniform(0.4, 1.6)))
        ok, r = ptref.refract(v, n, eta)
        cos_i = -(unit(v) @ n)
        sin_i = math.sqrt(max(0.0, 1 - cos_i ** 2))
        disc = 1 - eta ** 2 * (1 - cos_i ** 2)
        if abs(disc) < 1e-5:
            continue  # grazing the critical angle: either answer is within rounding
        assert ok == (disc > 0)
        if ok:
            r = r.astype(np.float64)
            assert abs(np.linalg.norm(r) - 1) < 1e-5
            sin_t = np.linalg.norm(np.cross(r, n))
            assert abs(sin_t - eta * sin_i) < 2e-5
            assert r @ n < 0
            # coplanar with v and n
            assert abs(np.cross(unit(v), n) @ r) < 2e-5


@pytest.mark.parametrize("nu,nv", [(0.1, 0.1), (0.3, 0.3), (0.7, 0.7), (0.2, 0.5)])
def test_ggx_normalisation(nu, nv):
    """Integral over the hemisphere (about +Y) of D(h) cos(theta_h) = 1."""
    a = max(nu, nv)
    s = np.linspace(0, 1, 241)[1:]
    # theta = atan(a * tan(pi/2 * s)) concentrates nodes where D is large
    st = np.tan(PI / 2 * (s - 0.5 / 240))
    theta = np.arctan(a * st)
    dtheta = a * (PI / 2) / np.cos(PI / 2 * (s - 0.5 / 240)) ** 2 / (1 + (a * st) ** 2) / 240
    phis = (np.arange(96) + 0.5) / 96 * 2 * PI
    total = 0.0
    for th, dth in zip(theta, dtheta):
        ring = 0.0
        for ph in phis:
            h = np.float32([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
            ring += ptref.ggx_d(h, nu, nv)
        total += ring * (2 * PI / 96) * math.cos(th) * math.sin(th) * dth
    assert abs(total - 1) < 5e-3, total


def test_power_heuristic():
    rng = np.random.default_rng(8)
    for a, b in rng.uniform(0.01, 100, (200, 2)).astype(np.float32):
        assert abs(ptref.power_heuristic(a, b) + ptref.power_heuristic(b, a) - 1) < 1e-6
    assert ptref.power_heuristic(3.0, 0.0) == 1.0
    assert ptref.power_heuristic(2.5, 2.5) == 0.5


# ------------------------------------------------------------- lights
@pytest.fixture(scope="module")
def light_scene():
    s = PyScene()
    m = s.lambert((1, 1, 1))
    s.spheres.append(((0, -1000, 0), 1.0, m))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    s.lights.append({"type": ptamd.PT_LIGHT_RECT, "axis": ptamd.PT_AXIS_Y, "rect": (213, 343, 227, 332, 554),
                     "emission": (15, 15, 15)})
    s.lights.append({"type": ptamd.PT_LIGHT_SPHERE, "center": (0, 10, 0), "radius": 2.0, "emission": (4, 4, 4)})
    orc = ptref.OracleScene(s.desc())
    yield orc
    orc.close()


def test_rect_light_samples_and_pdf(light_scene):
    orc = light_scene
    rng = np.random.default_rng(9)
    area = 130.0 * 105.0
    seed = 1234
    for _ in range(300):
        P = np.float32(rng.uniform([0, 0, 0], [555, 500, 555]))
        wi, _, seed = orc.light_sample(0, P, seed)
        q = P.astype(np.float64) + wi
        assert abs(q[1] - 554) < 1e-3
        assert 213 - 1e-3 <= q[0] <= 343 + 1e-3 and 227 - 1e-3 <= q[2] <= 332 + 1e-3
        pdf = orc.light_pdf(0, P, wi)
        d2 = wi.astype(np.float64) @ wi
        cos = abs(wi[1]) / math.sqrt(d2)
        assert abs(pdf - d2 / (cos * area)) <= 2e-5 * pdf
    P = np.float32([278, 100, 278])
    assert orc.light_pdf(0, P, np.float32([0, -1, 0])) == 0.0       # away from the light
    assert orc.light_pdf(0, P, np.float32([400, 454, 0])) == 0.0    # beside it


def test_sphere_light_pdf_closed_form(light_scene):
    orc = light_scene
    for P in (np.float32([0, 0, 0]), np.float32([3, -5, 1]), np.float32([0, 100, 0])):
        c = np.array([0, 10, 0.0])
        d2 = np.sum((c - P) ** 2)
        want = 1 / (2 * PI * (1 - math.sqrt(1 - 4.0 / d2)))
        toward = (c - P).astype(np.float32)
        got = orc.light_pdf(1, P, toward)
        assert abs(got - want) <= 1e-4 * want
        side = np.cross(toward, [1, 0, 0]).astype(np.float32)
        assert orc.light_pdf(1, P, side) == 0.0


# --------------------------------------------------------------- miss
def miss_scene(mtype, tex=None, spherical=1, c1=(1, 1, 1), c2=(0.5, 0.7, 1.0)):
    s = PyScene()
    m = s.lambert((1, 1, 1))
    s.spheres.append(((0, -1000, 0), 1.0, m))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    ti = -1
    if tex is not None:
        ti = s.texture(image=tex)
    s.miss = (mtype, c1, c2, ti, spherical)
    return ptref.OracleScene(s.desc())


def test_miss_gradient_and_constant():
    orc = miss_scene(ptamd.PT_MISS_GRADIENT)
    np.testing.assert_allclose(orc.miss([0, 1, 0]), [0.5, 0.7, 1.0], atol=1e-7)
    np.testing.assert_allclose(orc.miss([0, -3, 0]), [1, 1, 1], atol=1e-7)
    np.testing.assert_allclose(orc.miss([1, 0, 0]), [0.75, 0.85, 1.0], atol=1e-7)
    orc.close()
    orc = miss_scene(ptamd.PT_MISS_CONSTANT, c1=(0.25, 0.5, 2.0))
    np.testing.assert_array_equal(orc.miss([0.3, -0.2, 0.9]), np.float32([0.25, 0.5, 2.0]))
    orc.close()


def ramp_texture(w=8, h=4):
    """R = column index, G = row index: bilinear lookups return the
    continuous texel coordinate (u w - 1/2, v h - 1/2), wrapped."""
    img = np.zeros((h, w, 4), np.float32)
    img[:, :, 0] = np.arange(w)[None, :]
    img[:, :, 1] = np.arange(h)[:, None]
    return img


def test_miss_hdr_spherical_mapping():
    """u = atan2(z, x) / 2 pi, v = acos(y / |d|) / pi, times 2
    (miss.cu:64-95), with repeat wrap."""
    orc = miss_scene(ptamd.PT_MISS_HDR, tex=ramp_texture(), spherical=1)
    # direction -> (R, G) before the factor 2: x = u*8 - 0.5 wrapped blend
    cases = {(1, 0, 0): (3.5, 1.5), (0, 0, 1): (1.5, 1.5), (-1, 0, 0): (3.5, 1.5), (0, 0, -1): (5.5, 1.5),
             (0, 1, 0): (3.5, 1.5), (0, -1, 0): (3.5, 1.5), (0, 0.5, 0.5 * math.sqrt(3)): (1.5, 0.833333)}
    for d, (r, g) in cases.items():
        c = orc.miss(np.float32(d))
        assert abs(c[0] - 2 * r) < 2e-5 and abs(c[1] - 2 * g) < 2e-5, (d, c)
    orc.close()


def test_miss_hdr_cylindrical_and_image_mapping():
    orc = miss_scene(ptamd.PT_MISS_HDR, tex=ramp_texture(), spherical=0)
    # u = 1 - theta / 2pi with theta = atan2(x, z) in [0, 2pi); v = acos(y)/pi
    c = orc.miss(np.float32([0, 0, 1]))        # theta 0 -> u = 1 -> x = 7.5
    assert abs(c[0] - 2 * 3.5) < 2e-5 and abs(c[1] - 2 * 1.5) < 2e-5
    c = orc.miss(np.float32([1, 0, 0]))        # theta pi/2 -> u = 0.75 -> x = 5.5
    assert abs(c[0] - 2 * 5.5) < 2e-5
    c = orc.miss(np.float32([0, 1, 0]))        # v = 0 -> y = -0.5 -> rows 3, 0
    assert abs(c[1] - 2 * 1.5) < 2e-5
    orc.close()
    orc = miss_scene(ptamd.PT_MISS_IMAGE, tex=ramp_texture())
    # u = (atan2(x, z) + pi) / 2pi, v = (1 + sin(pi/2 - acos(y))) / 2, no factor
    c = orc.miss(np.float32([0, 0, 1]))        # u = 0.5 -> 3.5 ; v = 0.5 -> 1.5
    assert abs(c[0] - 3.5) < 2e-5 and abs(c[1] - 1.5) < 2e-5
    c = orc.miss(np.float32([-1, 0, 0]))       # u = 0.25 -> 1.5
    assert abs(c[0] - 1.5) < 2e-5
    c = orc.miss(np.float32([0, 0.6, 0.8]))    # v = 0.8 -> 2.7
    assert abs(c[1] - 2.7) < 2e-5
    orc.close()
