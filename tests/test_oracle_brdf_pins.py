"""Pins the oracle's Ashikhmin-Shirley, Torrance-Sparrow (GGX), Oren-Nayar
(sigma > 0) and metal-fuzz functions against what the mathematics fixes.

Samplers: chi-square tests of the oracle's samples against densities
derived here from the published models, in coordinates where the expected
distribution is uniform or integrated numerically:
  * Ashikhmin-Shirley half vectors (Ashikhmin & Shirley 2000, eq. 9):
    p(h) = sqrt((nu+1)(nv+1)) / (2 pi) cos(theta)^(nu cos^2 phi + nv sin^2 phi).
    Given phi, t = cos(theta)^(e(phi)+1) is uniform; the phi marginal
    integrates in closed form to atan(sqrt((nv+1)/(nu+1)) tan phi) / (2 pi)
    per quadrant.  The diffuse half is Malley's cosine sampling.
  * GGX visible normals (Heitz & d'Eon 2014): D_wo(h) = G1(wo) max(0, wo.h)
    D(h) / cos(theta_o), written in normalised slope space, where the
    density of (q = r^2/(1+r^2), psi) is G1/(2 pi wo_y) max(0, wo_y -
    r (wo_x ax cos psi + wo_z az sin psi)).
  * Metal fuzz (sampling.cuh:31-44): points uniform in the unit ball, so
    |p|^3 is uniform and directions are isotropic.
Evaluations: float64 re-derivations in angle / component form, with the
reference's conventions (angles against world +Y, the tangent frame
T = normalize(N x Y), B = T x N of ashikhmin_shirley.cuh:20-22).

Reference code followed: ashikhmin_shirley.cuh:11-151,
microfacets.cuh:11-20, 85-195, torrance_sparrow.cuh:10-64,
oren_nayar.cuh:22-65, materials.hpp:191-198 (Fujii's A, B), metal.cu:48-49,
sampling.cuh:31-44.
"""
import math

import numpy as np
import pytest
from scipy import stats

import ptamd
import ptref

PI = math.pi
Y = np.array([0.0, 1.0, 0.0])
AS, TS, ON = ptamd.PT_MAT_ASHIKHMIN_SHIRLEY, ptamd.PT_MAT_TORRANCE_SPARROW, ptamd.PT_MAT_OREN_NAYAR


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def f32unit(v):
    return unit(v).astype(np.float32)


def as_frame(n):
    """ashikhmin_shirley.cuh:20-22 / torrance_sparrow.cuh:18-20"""
    n = unit(n)
    t = unit(np.cross(n, Y))
    b = np.cross(t, n)
    return t, b, n


def half_vectors(wo, wi, n):
    """h = normalize(wo + wi) on the normal's side: a reflection about h and
    about -h is the same, so this recovers the sampled half vector."""
    h = unit(wo.astype(np.float64)[None, :] + wi.astype(np.float64))
    s = np.where(h @ unit(n) < 0, -1.0, 1.0)
    return h * s[:, None]


def chi2_pvalue(observed, expected, min_expected=5.0):
    observed = np.asarray(observed, np.float64).ravel()
    expected = np.asarray(expected, np.float64).ravel()
    keep = expected >= min_expected
    # samples where (almost) none are expected fail outright
    assert observed[~keep].sum() <= max(5.0, 3 * expected[~keep].sum() + 3 * math.sqrt(expected[~keep].sum())), \
        "samples in bins of negligible probability"
    o, e = observed[keep], expected[keep]
    e = e * o.sum() / e.sum()
    x2 = float(((o - e) ** 2 / e).sum())
    return stats.chi2.sf(x2, len(o) - 1), x2


# ------------------------------------------------------ Ashikhmin-Shirley
def as_phi_cdf(phi, nu, nv):
    """CDF of the half vector's azimuth under AS eq. 9 (phi from the B axis
    toward T): quadrant-wise atan(sqrt((nv+1)/(nu+1)) tan phi) / (2 pi)."""
    phi = np.mod(phi, 2 * PI)
    q = np.floor(phi / (PI / 2))
    r = phi - q * PI / 2
    k = math.sqrt((nv + 1) / (nu + 1))
    inner = np.arctan(k * np.tan(np.clip(r, 0, PI / 2 - 1e-12))) / (2 * PI)
    # quadrants 1 and 3 run mirrored (phi = pi - x, 2 pi - x)
    mirrored = (q == 1) | (q == 3)
    inner_m = 0.25 - np.arctan(k * np.tan(np.clip(PI / 2 - r, 0, PI / 2 - 1e-12))) / (2 * PI)
    return q * 0.25 + np.where(mirrored, inner_m, inner)


@pytest.mark.parametrize("nu,nv", [(100.0, 100.0), (10.0, 10.0), (10.0, 1000.0), (800.0, 30.0), (50.0, 200.0)])
def test_ashikhmin_shirley_half_vector_density(nu, nv):
    n = f32unit([0.3, 1.0, -0.4])
    wo = f32unit([0.2, 0.9, 0.5])
    wi, _, _, u0 = ptref.brdf_sample_many(AS, [0.5] * 3, [0.5] * 3, nu, nv, wo, n, 17, 120_000)
    spec = u0 >= 0.5
    assert abs(spec.mean() - 0.5) < 0.01
    h = half_vectors(wo, wi[spec], n)
    t, b, nn = as_frame(n)
    hx, hy, hz = h @ b, h @ nn, h @ t  # Spherical_Vector: (sin cos phi, cos, sin sin phi)
    phi = np.mod(np.arctan2(hz, hx), 2 * PI)
    e = nu * np.cos(phi) ** 2 + nv * np.sin(phi) ** 2
    tt = np.clip(hy, 0, 1) ** (e + 1)  # uniform given phi
    v = as_phi_cdf(phi, nu, nv)        # uniform
    hist, _, _ = np.histogram2d(tt, v, bins=[10, 12], range=[[0, 1], [0, 1]])
    pval, x2 = chi2_pvalue(hist, np.full(hist.shape, hist.sum() / hist.size))
    assert pval > 1e-4, (x2, pval)
    if nu != nv:
        # the same samples against the transposed lobe (nu on T) are rejected
        e_t = nv * np.cos(phi) ** 2 + nu * np.sin(phi) ** 2
        hist_t, _, _ = np.histogram2d(np.clip(hy, 0, 1) ** (e_t + 1), as_phi_cdf(phi, nv, nu), bins=[10, 12],
                                      range=[[0, 1], [0, 1]])
        assert chi2_pvalue(hist_t, np.full(hist.shape, hist.sum() / hist.size))[0] < 1e-12


def test_ashikhmin_shirley_diffuse_half_is_cosine_weighted():
    n = f32unit([-0.5, 0.7, 0.2])
    wo = f32unit([0.1, 0.8, -0.3])
    wi, _, _, u0 = ptref.brdf_sample_many(AS, [0.5] * 3, [0.5] * 3, 40.0, 400.0, wo, n, 5, 60_000)
    d = wi[u0 < 0.5].astype(np.float64)
    c = d @ unit(n)
    assert c.min() >= -1e-6
    # Malley: P(cos <= c) = c^2, azimuth uniform about N
    t, b, _ = as_frame(n)
    az = np.mod(np.arctan2(d @ b, d @ t), 2 * PI)
    hist, _, _ = np.histogram2d(np.clip(c, 0, 1) ** 2, az, bins=[10, 12], range=[[0, 1], [0, 2 * PI]])
    pval, x2 = chi2_pvalue(hist, np.full(hist.shape, hist.sum() / hist.size))
    assert pval > 1e-4, (x2, pval)


def as_eval64(rd, rs, nu, nv, wo, wi, n):
    """AS 2000 eqs. 4-5 in angle form with the reference's conventions:
    n.k of the Fresnel-like factors against world +Y (D9), the lobe's azimuth
    measured from T toward B in the evaluation (ashikhmin_shirley.cuh:106),
    and the value returned as rho_d / p_d + rho_s / p_s with p_s the
    reference's |h.k| / (norm * lobe) (:145-151)."""
    rd, rs = np.asarray(rd, np.float64), np.asarray(rs, np.float64)
    t, b, nn = as_frame(n)
    wo, wi = unit(wo), unit(wi)
    ci, co = abs(wi[1]), abs(wo[1])
    rho_d = 28.0 / (23.0 * PI) * rd * (1 - rs) * (1 - (1 - ci / 2) ** 5) * (1 - (1 - co / 2) ** 5)
    p_d = wi @ nn / PI
    if p_d < 0:
        return np.zeros(3), 1.0
    h = unit(wo + wi)
    hk = abs(h @ wi)
    theta = math.acos(min(1.0, abs(h @ nn)))
    phi = math.atan2(abs(h @ b), abs(h @ t))
    expo = nu if nu == nv else nu * math.cos(phi) ** 2 + nv * math.sin(phi) ** 2
    norm = math.sqrt((nu + 1) * (nv + 1)) / (8 * PI)
    lobe = math.cos(theta) ** expo
    fres = rs + (1 - rs) * (1 - hk) ** 5
    rho_s = norm * lobe * fres / (hk * max(ci, co))
    p_s = hk / (norm * lobe)
    return rho_d / p_d + rho_s / p_s, (p_d + p_s) / 2


def test_ashikhmin_shirley_eval_matches_float64():
    rng = np.random.default_rng(31)
    checked = 0
    for k in range(600):
        n = f32unit(rng.normal(size=3) + [0, 2.5, 0])
        t, b, nn = as_frame(n)
        nu, nv = (float(x) for x in np.float32(10 ** rng.uniform(1, 3, 2)))
        if k % 3 == 0:
            nv = nu
        rd = np.float32(rng.uniform(0.05, 0.95, 3))
        rs = np.float32(rng.uniform(0.05, 0.95, 3))
        wo = f32unit(rng.normal(size=3) * [1, 0.3, 1] + nn * 1.5)
        # a half vector near N, so that the lobe is not negligible
        hh = unit(nn + rng.normal(size=3) * 0.08)
        wi = f32unit(-wo + 2 * (wo @ hh) * hh)
        if wi.astype(np.float64) @ nn < 0.05 or wo.astype(np.float64) @ nn < 0.05:
            continue
        f, pdf = ptref.brdf_eval(AS, rd, rs, nu, nv, wo, wi, n)
        ef, epdf = as_eval64(rd, rs, nu, nv, wo.astype(np.float64), wi.astype(np.float64), n.astype(np.float64))
        # powf of exponents up to 1000 amplifies float32 rounding of |h.n|
        np.testing.assert_allclose(f, ef, rtol=3e-3)
        assert abs(pdf - epdf) <= 3e-3 * abs(epdf)
        checked += 1
    assert checked > 300
    # below the surface: zero, pdf untouched
    n = f32unit([0, 1, 0.2])
    f, pdf = ptref.brdf_eval(AS, [0.5] * 3, [0.5] * 3, 10.0, 20.0, f32unit([0.1, 1, 0]), f32unit([0.3, -1, 0]), n)
    assert np.all(f == 0) and pdf == 1.0


def test_ashikhmin_shirley_mirror_direction_closed_form():
    """h = N (wi the mirror of wo): lobe = 1 for any nu, nv, and with Rd = 0,
    Rs = 1 the value is norm^2 / ((wo.N)^2 max(|wi_y|, |wo_y|))."""
    rng = np.random.default_rng(2)
    for _ in range(100):
        n = f32unit(rng.normal(size=3) + [0, 2, 0])
        nn = n.astype(np.float64)
        wo = f32unit(rng.normal(size=3) * 0.5 + nn)
        c = float(wo.astype(np.float64) @ nn)
        if c < 0.1:
            continue
        wi = f32unit(-wo + 2 * c * nn)
        nu, nv = float(rng.uniform(10, 500)), float(rng.uniform(10, 500))
        f, _ = ptref.brdf_eval(AS, [0, 0, 0], [1, 1, 1], np.float32(nu), np.float32(nv), wo, wi, n)
        norm = math.sqrt((np.float32(nu) + 1.0) * (np.float32(nv) + 1.0)) / (8 * PI)
        expect = norm ** 2 / (c * c * max(abs(float(wi[1])), abs(float(wo[1]))))
        np.testing.assert_allclose(f, expect, rtol=2e-3)


# ------------------------------------------------------- Torrance-Sparrow
def smith_lambda64(w, ax, az):
    """Heitz 2014 eq. 72 in component form (y up)."""
    return (-1 + math.sqrt(1 + (ax * ax * w[0] ** 2 + az * az * w[2] ** 2) / (w[1] ** 2))) / 2


def ggx_d64(h, ax, az):
    if h[1] <= 0:
        return 0.0
    return 1.0 / (PI * ax * az * h[1] ** 4 * (1 + (h[0] ** 2 / ax ** 2 + h[2] ** 2 / az ** 2) / h[1] ** 2) ** 2)


@pytest.mark.parametrize("ax,az,wo", [
    (0.3, 0.3, [0.2, 0.9, 0.5]),
    (0.1, 0.5, [0.6, 0.5, -0.3]),
    (0.5, 0.02, [-0.4, 0.35, 0.8]),
    (0.04, 0.04, [0.0, 1.0, 0.0]),   # normal incidence: the sampler's special branch
    (0.36, 0.09, [0.9, 0.2, 0.1]),   # grazing
])
def test_torrance_sparrow_samples_visible_normals(ax, az, wo):
    n = f32unit([0.25, 1.0, 0.1])
    wo = f32unit(wo)
    ax, az = float(np.float32(ax)), float(np.float32(az))
    wi, _, _, _ = ptref.brdf_sample_many(TS, [0.9] * 3, [0] * 3, ax, az, wo, n, 3, 100_000)
    h = half_vectors(wo, wi, n)
    t, b, nn = as_frame(n)
    hl = np.stack([h @ b, h @ nn, h @ t], axis=1)  # the sampler's local frame (y up)
    assert hl[:, 1].min() > 0
    u = (-hl[:, 0] / hl[:, 1]) / ax
    v = (-hl[:, 2] / hl[:, 1]) / az
    r2 = u * u + v * v
    q = r2 / (1 + r2)
    psi = np.mod(np.arctan2(v, u), 2 * PI)
    nq, npsi = 12, 12
    hist, _, _ = np.histogram2d(q, psi, bins=[nq, npsi], range=[[0, 1], [0, 2 * PI]])
    # expected mass per bin: the density of (q, psi), the wo of the sampler
    # being the world direction read in the local frame
    w = wo.astype(np.float64)
    g1 = 1 / (1 + smith_lambda64(w, ax, az))
    # exact in q (the antiderivative of r(q) = sqrt(q / (1 - q)) is
    # asin(sqrt q) - sqrt(q (1 - q))), midpoint rule in psi
    m = 400
    pp = (np.arange(npsi * m) + 0.5) / (npsi * m) * 2 * PI
    c = w[0] * ax * np.cos(pp) + w[2] * az * np.sin(pp)
    qcut = np.where(c > 0, (w[1] / np.where(c > 0, c, 1)) ** 2 / (1 + (w[1] / np.where(c > 0, c, 1)) ** 2), 1.0)

    def rint(qv):
        return np.arcsin(np.sqrt(qv)) - np.sqrt(qv * (1 - qv))

    edges = np.linspace(0, 1, nq + 1)
    expected = np.zeros((nq, npsi))
    for i in range(nq):
        q0 = np.full_like(pp, edges[i])
        q1 = np.minimum(edges[i + 1], qcut)
        ok = q1 > q0
        mass = np.where(ok, w[1] * (q1 - q0) - c * (rint(np.where(ok, q1, q0)) - rint(q0)), 0.0)
        expected[i] = (g1 / (2 * PI * w[1]) * mass * (2 * PI / (npsi * m))).reshape(npsi, m).sum(axis=1)
    # Smith's G1 normalises the visible-normal density
    assert abs(expected.sum() - 1) < 2e-3
    pval, x2 = chi2_pvalue(hist, expected * hist.sum())
    assert pval > 1e-4, (x2, pval)


def ts_eval64(rs, ax, az, wo, wi):
    rs = np.asarray(rs, np.float64)
    wo, wi = unit(wo), unit(wi)
    ci, co = max(wi[1], 1e-6), max(wo[1], 1e-6)
    h = unit(wo + wi)
    hk = abs(h @ wi)
    fres = rs + (1 - rs) * (1 - hk) ** 5
    d = ggx_d64(h, ax, az)
    g = 1 / (1 + smith_lambda64(wo, ax, az) + smith_lambda64(wi, ax, az))
    return rs * d * g * fres / (4 * ci * co), d * abs(h[1]) / (4 * (wo @ h))


def test_torrance_sparrow_eval_matches_float64():
    rng = np.random.default_rng(8)
    for k in range(500):
        ax, az = (float(x) for x in np.float32(rng.uniform(0.05, 0.6, 2) ** 2))
        if k % 3 == 0:
            az = ax
        rs = np.float32(rng.uniform(0.05, 1, 3))
        wo = f32unit(rng.normal(size=3) + [0, 1.2, 0])
        wi = f32unit(rng.normal(size=3) + [0, 1.2, 0])
        if wo[1] < 0.05 or wi[1] < 0.05:
            continue
        n = f32unit(rng.normal(size=3) + [0, 3, 0])  # evaluation is against world +Y (D9)
        f, pdf = ptref.brdf_eval(TS, rs, [0] * 3, ax, az, wo, wi, n)
        ef, epdf = ts_eval64(rs, ax, az, wo.astype(np.float64), wi.astype(np.float64))
        np.testing.assert_allclose(f, ef, rtol=2e-4, atol=1e-30)
        assert abs(pdf - epdf) <= 2e-4 * abs(epdf) + 1e-30
        # reciprocity: D(h), F(h.wi = h.wo) and the Smith G are symmetric
        f2, _ = ptref.brdf_eval(TS, rs, [0] * 3, ax, az, wi, wo, n)
        np.testing.assert_allclose(f2, f, rtol=2e-6)


def test_torrance_sparrow_normal_incidence_closed_form():
    """wo = wi = +Y: Lambda(0) = 0 so G = 1, D(+Y) = 1/(pi ax az), F = Rs:
    f = Rs^2 / (4 pi ax az); pdf = D / 4.  The float32 D loses
    ~2^-24 / (ax az) relative to cancellation in (a^2 - 1) cos^2 + 1, as
    the reference's formula does (microfacets.cuh:89-91)."""
    for ax, az in [(0.01, 0.01), (0.2, 0.05), (0.36, 0.36)]:
        ax, az = float(np.float32(ax)), float(np.float32(az))
        rtol = 1e-5 + 4 * 2.0 ** -24 / (ax * az)
        rs = np.float32([0.9, 0.5, 0.2])
        f, pdf = ptref.brdf_eval(TS, rs, [0] * 3, ax, az, f32unit([0, 1, 0]), f32unit([0, 1, 0]),
                                 f32unit([0.1, 1, 0]))
        np.testing.assert_allclose(f, rs.astype(np.float64) ** 2 / (4 * PI * ax * az), rtol=rtol)
        assert abs(pdf - 1 / (4 * PI * ax * az)) <= rtol * pdf


# --------------------------------------------------------------- Oren-Nayar
def fujii_ab(sigma):
    """Fujii's improved Oren-Nayar (materials.hpp:191-198)."""
    a = 1.0 / (PI + (PI / 2 - 2.0 / 3.0) * sigma)
    return a, sigma * a


def on_eval64(rho, a, b, wo, wi, n):
    """Qualitative ON in angle form against world +Y (D9), times the
    reference's 1/pi (D10); pdf = cos(wi, N) / pi."""
    rho = np.asarray(rho, np.float64)
    wo, wi = unit(wo), unit(wi)
    ti, to = math.acos(min(1.0, abs(wi[1]))), math.acos(min(1.0, abs(wo[1])))
    if math.sin(ti) > 1e-4 and math.sin(to) > 1e-4:
        dphi = math.atan2(wi[2], wi[0]) - math.atan2(wo[2], wo[0])
        mc = max(0.0, math.cos(dphi))
    else:
        mc = 0.0
    alpha, beta = max(ti, to), min(ti, to)
    return rho / PI * (a + b * mc * math.sin(alpha) * math.tan(beta)), wi @ unit(n) / PI


def test_oren_nayar_eval_matches_float64():
    rng = np.random.default_rng(12)
    for _ in range(500):
        sigma = float(rng.uniform(0, 1))
        a, b = (float(x) for x in np.float32(fujii_ab(sigma)))
        rho = np.float32(rng.uniform(0.05, 0.95, 3))
        wo = f32unit(rng.normal(size=3) + [0, 0.8, 0])
        wi = f32unit(rng.normal(size=3) + [0, 0.8, 0])
        if abs(wo[1]) < 0.05 or abs(wi[1]) < 0.05:
            continue
        n = f32unit(rng.normal(size=3) + [0, 2, 0])
        f, pdf = ptref.brdf_eval(ON, rho, rho, a, b, wo, wi, n)
        ef, epdf = on_eval64(rho, a, b, wo.astype(np.float64), wi.astype(np.float64), n.astype(np.float64))
        np.testing.assert_allclose(f, ef, rtol=2e-5)
        assert abs(pdf - epdf) <= 2e-6 * abs(epdf) + 1e-7


def test_oren_nayar_backscatter_exceeds_lambert():
    """Retro-reflection: for wi = wo off the normal the ON term is larger
    than A, and it equals A when the azimuths are opposite."""
    a, b = (float(x) for x in np.float32(fujii_ab(0.8)))
    rho = np.float32([1, 1, 1])
    n = f32unit([0, 1, 0.3])
    w = f32unit([0.6, 0.5, 0.2])
    f_back, _ = ptref.brdf_eval(ON, rho, rho, a, b, w, w, n)
    opp = f32unit([-0.6, 0.5, -0.2])
    f_opp, _ = ptref.brdf_eval(ON, rho, rho, a, b, w, opp, n)
    th = math.acos(float(w[1]))
    np.testing.assert_allclose(f_back, (a + b * math.sin(th) * math.tan(th)) / PI, rtol=2e-5)
    np.testing.assert_allclose(f_opp, a / PI, rtol=2e-6)


def test_oren_nayar_host_parameters_are_fujii():
    """The mesh scene's Oren-Nayar materials (sigma drawn in [0, 30 deg])
    carry A, B with B / A = sigma and 1 / A = pi + (pi/2 - 2/3) sigma."""
    sc = ptamd.Scene(ptamd.PT_SCENE_MESHES, 64, 36, 3)
    seen = 0
    for i in range(sc.desc.num_materials):
        m = sc.desc.materials[i]
        if m.type != ON:
            continue
        a, b = float(m.param[0]), float(m.param[1])
        sigma = b / a
        assert 0 <= sigma <= math.radians(30) + 1e-6
        assert abs(1 / a - (PI + (PI / 2 - 2 / 3) * sigma)) < 2e-6 * (1 / a)
        seen += 1
    assert seen >= 2


# ------------------------------------------------------------- metal fuzz
def test_metal_fuzz_is_uniform_in_the_ball():
    n = f32unit([0.2, 1.0, -0.1])
    wo = f32unit([0.5, 0.7, 0.3])
    fuzz = np.float32(0.7)
    d = ptref.metal_scatter_many(wo, n, float(fuzz), 9, 80_000).astype(np.float64)
    refl = (-wo).astype(np.float64) - 2 * ((-wo).astype(np.float64) @ n.astype(np.float64)) * n.astype(np.float64)
    p = (d - refl) / float(fuzz)
    r = np.linalg.norm(p, axis=1)
    assert r.max() <= 1 + 1e-5
    # radius: P(|p| <= s) = s^3
    ks = stats.kstest(np.clip(r, 0, 1) ** 3, "uniform")
    assert ks.pvalue > 1e-4, ks
    # direction: cos(polar) uniform in [-1, 1], azimuth uniform
    u = p / r[:, None]
    hist, _, _ = np.histogram2d(u[:, 2], np.mod(np.arctan2(u[:, 1], u[:, 0]), 2 * PI), bins=[10, 12],
                                range=[[-1, 1], [0, 2 * PI]])
    pval, x2 = chi2_pvalue(hist, np.full(hist.shape, hist.sum() / hist.size))
    assert pval > 1e-4, (x2, pval)
    # the ball is centred on the mirror direction
    assert np.linalg.norm(p.mean(axis=0)) < 0.01
