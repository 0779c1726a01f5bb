"""Pins the oracle's environment importance sampling (no reference
counterpart: the reference only looks the HDR map up on a miss,
programs/miss.cu:64-95) and the environment MIS of the integrator.

  * tables: inclusive prefix sums of the 64-bit fixed-point texel weights
    (luminance x sin(theta), scaled by 2^(40 - e)) against numpy cumsum;
  * sampling: texel frequencies against the table probabilities (chi^2),
    the returned pdf against p(u, v) W H / (2 pi^2 sin(theta)), and the pdf
    integrating to 1 over the sphere;
  * integrator: a white Lambertian sphere under a constant environment with
    environment NEE + MIS has expectation exactly 1 per pixel.
"""
import math

import numpy as np
import pytest

import ptamd
import ptref
from helpers import PyScene
from test_oracle_integrator import camera


def env_scene(img, with_light=True, sphere=True):
    s = PyScene()
    m = s.lambert((1, 1, 1))
    if sphere:
        s.spheres.append(((0, 0, 0), 1.0, m))
    else:
        s.spheres.append(((0, -1e4, 0), 1.0, m))  # out of the way
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    t = s.texture(image=img)
    s.miss = (ptamd.PT_MISS_HDR, (0, 0, 0), (0, 0, 0), t, 1)
    if with_light:
        s.lights.append({"type": ptamd.PT_LIGHT_ENV, "emission": (1, 1, 1)})
    return s


def sky(W=64, H=32, seed=0):
    rng = np.random.default_rng(seed)
    img = np.zeros((H, W, 4), np.float32)
    img[..., :3] = rng.uniform(0, 1, (H, W, 3)) ** 4  # strongly varying
    img[H // 3, W // 5, :3] = 500.0                    # a hot texel
    img[5, 7, :3] = -1.0                               # negative: no weight
    img[6, 9, 0] = np.nan                              # NaN: no weight
    return img


def reference_tables(img):
    H, W = img.shape[:2]
    t = img.astype(np.float64)
    lum = 0.2126 * t[..., 0] + 0.7152 * t[..., 1] + 0.0722 * t[..., 2]
    st = ptref.libm("sin", (np.float32(math.pi) * ((np.arange(H, dtype=np.float32) + np.float32(0.5)) /
                                                   np.float32(H))).astype(np.float32)).astype(np.float64)
    w = lum * st[:, None]
    w = np.where(w > 0, w, 0.0)  # NaN compares false
    _, ex = np.frexp(w.max())
    q = np.floor(w * np.ldexp(1.0, 40 - int(ex))).astype(np.uint64)
    cond = np.cumsum(q, axis=1, dtype=np.uint64)
    marg = np.cumsum(cond[:, -1], dtype=np.uint64)
    return q, cond, marg


def test_tables_are_fixed_point_prefix_sums():
    img = sky()
    orc = ptref.OracleScene(env_scene(img).desc())
    cond, marg, total = orc.env_tables()
    q, rc, rm = reference_tables(img)
    np.testing.assert_array_equal(cond, rc)
    np.testing.assert_array_equal(marg, rm)
    assert total == int(rm[-1]) and q.max() < 2 ** 40 <= 2 * q.max()
    assert q[5, 7] == 0 and q[6, 9] == 0
    orc.close()


def test_no_tables_without_environment_light():
    orc = ptref.OracleScene(env_scene(sky(), with_light=False).desc())
    assert orc.env_tables()[2] == 0
    orc.close()


def test_environment_light_requires_spherical_hdr():
    s = env_scene(sky())
    s.miss = (ptamd.PT_MISS_HDR, (0, 0, 0), (0, 0, 0), 0, 0)  # cylindrical mapping
    with pytest.raises(RuntimeError):
        ptref.OracleScene(s.desc())


def dir_to_texel(d, W, H):
    d = d.astype(np.float64)
    u = np.arctan2(d[:, 2], d[:, 0]) / (2 * math.pi)
    v = np.arccos(np.clip(d[:, 1] / np.linalg.norm(d, axis=1), -1, 1)) / math.pi
    i = np.clip(np.floor((u - np.floor(u)) * W), 0, W - 1).astype(int)
    j = np.clip(np.floor(v * H), 0, H - 1).astype(int)
    return i, j


def test_sampling_follows_the_table_and_pdf_closed_form():
    img = sky(32, 16, seed=3)
    H, W = img.shape[:2]
    orc = ptref.OracleScene(env_scene(img).desc())
    q, _, _ = reference_tables(img)
    prob = q / q.sum()
    n = 60000
    seed = 77
    dirs = np.zeros((n, 3), np.float32)
    pdfs = np.zeros(n)
    P = np.float32([0, 5, 0])
    for k in range(n):
        wi, pdf, seed = orc.light_sample(0, P, seed)
        dirs[k] = wi
        pdfs[k] = pdf
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1, atol=1e-5)
    i, j = dir_to_texel(dirs, W, H)
    counts = np.zeros((H, W))
    np.add.at(counts, (j, i), 1)
    assert counts[prob == 0].sum() <= 2  # only boundary rounding may land there
    exp = prob * n
    keep = exp > 5
    chi2 = np.sum((counts[keep] - exp[keep]) ** 2 / exp[keep])
    dof = keep.sum() - 1
    assert chi2 < dof + 5 * math.sqrt(2 * dof), (chi2, dof)
    # pdf = p(texel) W H / (2 pi^2 sin(theta))
    sin_t = np.sqrt(np.maximum(0, 1 - dirs[:, 1].astype(np.float64) ** 2))
    want = prob[j, i] * W * H / (2 * math.pi ** 2 * sin_t)
    ok = (sin_t > 1e-3) & (prob[j, i] > 0)
    rel = np.abs(pdfs[ok] - want[ok]) / want[ok]
    assert np.median(rel) < 1e-4 and np.mean(rel < 1e-3) > 0.99
    orc.close()


def test_pdf_integrates_to_one():
    img = sky(32, 16, seed=4)
    orc = ptref.OracleScene(env_scene(img).desc())
    nt, npf = 192, 384
    total = 0.0
    P = np.float32([0, 5, 0])
    for a in range(nt):
        th = (a + 0.5) / nt * math.pi
        row = 0.0
        for b in range(npf):
            ph = (b + 0.5) / npf * 2 * math.pi
            d = np.float32([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
            row += orc.light_pdf(0, P, d)
        total += row * math.sin(th) * (math.pi / nt) * (2 * math.pi / npf)
    assert abs(total - 1) < 5e-3, total
    orc.close()


def test_env_mis_furnace():
    """rho = 1 Lambertian sphere under a constant environment of radiance 1
    (texture 0.5, x2): light-sample and BRDF-sample estimates carry power-
    heuristic weights that sum to one, so each pixel has expectation 1."""
    img = np.full((32, 64, 4), 0.5, np.float32)
    s = env_scene(img)
    orc = ptref.OracleScene(s.desc())
    W = H = 24
    p = ptamd.RenderParams(W, H, 6, 0, 11, 0, ptamd.PT_FLAG_ENV_MIS, 0)
    acc, (rgb, sig, b), st = orc.render(camera((0, 0, 3.5), (0, 0, 0), 40, 1), p, 0, 256, keep_last=True)
    img_mean = acc / 256
    hit = b.reshape(H, W) > 1
    assert hit.mean() > 0.3
    assert abs(img_mean[hit].mean() - 1) < 0.01
    assert abs(img_mean.mean() - 1) < 0.01
    # pixels that never met the sphere see the unclamped sky, exactly 1
    assert np.mean(np.all(img_mean == 1.0, axis=-1)) > 0.3
    assert int(st[2]) > 0                                  # environment shadow rays were traced
    orc.close()


def test_env_mis_off_keeps_reference_miss():
    """Without PT_FLAG_ENV_MIS the environment light is ignored and a miss
    returns L + clamp(2 tex T, 0, 1) (raygen.cu:81-82)."""
    img = np.full((32, 64, 4), 3.0, np.float32)
    orc = ptref.OracleScene(env_scene(img, sphere=False).desc())
    p = ptamd.RenderParams(8, 8, 4, 0, 11, 0, 0, 0)
    acc, _, st = orc.render(camera((0, 0, 3.5), (0, 0, 0), 40, 1), p, 0, 1)
    np.testing.assert_array_equal(acc, 1.0)
    assert int(st[2]) == 0
    p.flags = ptamd.PT_FLAG_ENV_MIS
    acc, _, _ = orc.render(camera((0, 0, 3.5), (0, 0, 0), 40, 1), p, 0, 1)
    np.testing.assert_array_equal(acc, 6.0)
    orc.close()
