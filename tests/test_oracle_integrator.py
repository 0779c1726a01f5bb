"""Pins the oracle's integrator (programs/raygen.cu:66-168 color() and
renderPixel, diffuse_light.cu:30-48, light_sample.cuh:35-103) with scenes
whose answer is known exactly under the reference's semantics."""
import math

import numpy as np
import pytest

import ptamd
import ptref
from helpers import PyScene


def camera(lookfrom, lookat, vfov, aspect, vup=(0, 1, 0)):
    """Pinhole camera of host_includes/camera.hpp:13-51."""
    o = np.array(lookfrom, np.float64)
    w = o - np.array(lookat, np.float64)
    w /= np.linalg.norm(w)
    u = np.cross(vup, w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    hh = math.tan(math.radians(vfov) / 2)
    hw = aspect * hh
    c = ptamd.Camera()
    for name, val in (("lower_left", o - hw * u - hh * v - w), ("horizontal", 2 * hw * u),
                      ("vertical", 2 * hh * v), ("origin", o), ("u", u), ("v", v), ("w", w)):
        getattr(c, name)[:] = [float(x) for x in np.float32(val)]
    c.lens_radius, c.time0, c.time1 = 0.0, 0.0, 1.0
    return c


def params(W, H, depth, russian=0, rr_min=11):
    return ptamd.RenderParams(W, H, depth, russian, rr_min, 0, 0, 0)


def test_white_furnace():
    """A white Lambertian sphere (rho = 1) under a constant white sky, no
    lights, no roulette: f / pdf = 1 at every bounce and every path escapes
    after its first bounce, so every path returns exactly 1."""
    s = PyScene()
    s.spheres.append(((0, 0, 0), 1.0, s.lambert((1, 1, 1))))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    s.miss = (ptamd.PT_MISS_CONSTANT, (1, 1, 1), (1, 1, 1), -1, 1)
    orc = ptref.OracleScene(s.desc())
    W = H = 32
    acc, (rgb, sig, bounces), st = orc.render(camera((0, 0, 4), (0, 0, 0), 40, 1), params(W, H, 8), 0, 4,
                                              keep_last=True)
    # f / pdf is 1 up to the rounding of one reciprocal multiply (a vector
    # divided by a scalar is a reciprocal and three products)
    exact = np.all(rgb == 1.0, axis=1)
    assert exact.mean() > 0.85
    assert np.all(np.abs(rgb - 1) < 1e-6)
    np.testing.assert_allclose(acc / 4, 1.0, atol=1e-6)
    assert set(np.unique(bounces)) <= {1, 2}
    assert (bounces == 2).sum() > 100  # the sphere is in view
    orc.close()


def closed_box(s, mat, lo=-2.0, hi=2.0):
    """Six rects enclosing the cube [lo, hi]^3."""
    for axis in (0, 1, 2):
        for k, flip in ((lo, 0), (hi, 1)):
            s.rects.append((axis, flip, lo, hi, lo, hi, k, mat))
    s.instance(ptamd.PT_GEOM_RECTS, 0, 6)


def test_closed_box_without_light_is_black():
    """No light and no roulette: every path exhausts max_depth and returns 0
    (raygen.cu:121-122), whatever the (bright) sky."""
    s = PyScene()
    closed_box(s, s.lambert((0.9, 0.9, 0.9)))
    s.miss = (ptamd.PT_MISS_CONSTANT, (5, 5, 5), (5, 5, 5), -1, 1)
    orc = ptref.OracleScene(s.desc())
    acc, (rgb, sig, bounces), st = orc.render(camera((0, 0, 1), (0, 0, 0), 60, 1), params(16, 16, 6), 0, 2,
                                              keep_last=True)
    assert np.all(acc == 0.0)
    assert np.all(bounces == 6)
    assert int(st[1]) == 16 * 16 * 2 * 6
    orc.close()


def test_closed_box_roulette_returns_L_plus_T():
    """With roulette, a terminated path returns L + T (raygen.cu:110-118):
    in a dark closed box L = 0, so the result equals the surviving
    throughput, which is bounded by rho^depth / prob products > 0."""
    s = PyScene()
    closed_box(s, s.lambert((0.5, 0.5, 0.5)))
    orc = ptref.OracleScene(s.desc())
    acc, (rgb, sig, bounces), _ = orc.render(camera((0, 0, 1), (0, 0, 0), 60, 1), params(16, 16, 50, 1, 2), 0, 1,
                                             keep_last=True)
    ended = bounces < 50
    assert ended.mean() > 0.99
    assert np.all(rgb[ended] > 0)
    assert np.all(rgb[~ended] == 0)
    orc.close()


@pytest.mark.parametrize("flip", [0, 1])
def test_emitter_seen_directly(flip):
    """Camera rays hitting the only light: Direct_Light returns 0 after one
    draw (single light), T *= Le only when dot(N, Wo) < 0
    (diffuse_light.cu:43-47), and depth 0 adds T: the pixel is Le or 1."""
    s = PyScene()
    le = (4.0, 5.0, 6.0)
    lm = s.material(ptamd.PT_MAT_DIFFUSE_LIGHT, s.texture(le))
    s.rects.append((ptamd.PT_AXIS_Z, flip, -10, 10, -10, 10, 0.0, lm))
    s.instance(ptamd.PT_GEOM_RECTS, 0, 1)
    s.lights.append({"type": ptamd.PT_LIGHT_RECT, "axis": ptamd.PT_AXIS_Z, "rect": (-10, 10, -10, 10, 0.0),
                     "emission": le})
    s.miss = (ptamd.PT_MISS_CONSTANT, (0, 0, 0), (0, 0, 0), -1, 1)
    orc = ptref.OracleScene(s.desc())
    _, (rgb, sig, bounces), st = orc.render(camera((0, 0, 5), (0, 0, 0), 30, 1), params(8, 8, 4), 0, 1,
                                            keep_last=True)
    # normal +Z (flip 0) faces the camera: dot(N, Wo) > 0, T stays 1
    want = np.float32([1, 1, 1]) if flip == 0 else np.float32(le)
    np.testing.assert_array_equal(rgb, np.broadcast_to(want, rgb.shape))
    assert np.all(bounces == 1)
    assert int(st[2]) == 0  # no shadow ray: single light seen by an emitter
    orc.close()


def lit_floor_scene():
    s = PyScene()
    s.rects.append((ptamd.PT_AXIS_Y, 0, -50, 50, -50, 50, 0.0, s.lambert((0.5, 0.5, 0.5))))
    lm = s.material(ptamd.PT_MAT_DIFFUSE_LIGHT, s.texture((10, 10, 10)))
    s.rects.append((ptamd.PT_AXIS_Y, 1, -1, 1, -1, 1, 5.0, lm))
    s.instance(ptamd.PT_GEOM_RECTS, 0, 2)
    s.lights.append({"type": ptamd.PT_LIGHT_RECT, "axis": ptamd.PT_AXIS_Y, "rect": (-1, 1, -1, 1, 5.0),
                     "emission": (10, 10, 10)})
    s.miss = (ptamd.PT_MISS_CONSTANT, (0, 0, 0), (0, 0, 0), -1, 1)
    return s


def test_direct_lighting_of_a_floor_converges_to_closed_form():
    """A Lambertian floor under a small square light: the MIS estimate of
    direct light at the first hit approaches rho/pi * Le * integral(cos_i
    cos_l / d^2 dA) at the point below the light.  Max depth 2: the second
    segment only ends the path (light seen at depth 1 adds nothing, a miss
    adds the black sky), while at max depth 1 exhaustion would discard L."""
    s = lit_floor_scene()
    orc = ptref.OracleScene(s.desc())
    cam = camera((0, 3, 0.001), (0, 0, 0), 1.0, 1)  # a narrow view of the point under the light
    acc, _, _ = orc.render(cam, params(4, 4, 2), 0, 4096, threads=8)
    mean = acc.mean(axis=(0, 1)) / 4096
    # exact: E = Le * integral over the square of h^2 / (h^2 + x^2 + z^2)^2
    h, n = 5.0, 400
    g = (np.arange(n) + 0.5) / n * 2 - 1
    X, Z = np.meshgrid(g, g)
    E = 10 * np.sum(h * h / (h * h + X ** 2 + Z ** 2) ** 2) * (2 / n) ** 2
    want = 0.5 / math.pi * E
    assert abs(mean[0] - want) < 0.02 * want, (mean, want)
    orc.close()


def test_render_is_deterministic_and_splits():
    s = lit_floor_scene()
    orc = ptref.OracleScene(s.desc())
    cam = camera((0, 3, 6), (0, 0, 0), 50, 1)
    p = params(24, 24, 5, 1, 2)
    a1, (r1, g1, b1), _ = orc.render(cam, p, 0, 6, threads=1, keep_last=True)
    a8, (r8, g8, b8), _ = orc.render(cam, p, 0, 6, threads=8, keep_last=True)
    assert np.array_equal(a1, a8) and np.array_equal(r1, r8) and np.array_equal(g1, g8)
    part, _, _ = orc.render(cam, p, 0, 2, threads=4)
    part, _, _ = orc.render(cam, p, 2, 4, threads=4, accum=part)
    assert np.array_equal(part, a1)
    orc.close()


def test_display_tone_map():
    """make_Color (raygen.cu:136-146): int(255.99 * clamp(sqrt(avg)))."""
    acc = np.array([[[0.0, 0.25, 4.0], [1.0, 0.01, 0.5]]], np.float64) * 4
    out = ptref.display(acc, 4)
    assert out.shape == (1, 2, 4)
    np.testing.assert_array_equal(out[0, 0, :3], [0, int(255.99 * 0.5), 255])
    np.testing.assert_array_equal(out[0, 1, :3], [255, int(255.99 * np.float32(0.1)), int(255.99 * np.sqrt(np.float32(0.5)))])
    assert np.all(out[..., 3] == 255)


def test_nan_paths_are_counted_and_removed():
    """raygen.cu:126-134 de_nan zeroes NaN channels before accumulation; the
    render statistics count those paths.  An Ashikhmin-Shirley floor facing
    +Y has a NaN tangent frame (D9): its continuation is NaN and, under an
    MIS environment (unclamped), so is the path.  The Cornell box has none."""
    from test_gpu_fullsize import nan_scene

    s = nan_scene()
    W, H = 32, 24
    prm = ptamd.RenderParams(W, H, 8, 1, 3, 0, ptamd.PT_FLAG_ENV_MIS, 0)
    acc, (rgb, _, _), st = ptref.OracleScene(s.desc()).render(camera((0, 3, 8), (0, 0, 0), 40, W / H), prm, 0,
                                                                2, keep_last=True)
    nan_last = int(np.any(np.isnan(rgb), axis=1).sum())
    assert 0 < nan_last <= int(st[3])
    assert np.all(np.isfinite(acc))
    sc = ptamd.Scene(ptamd.PT_SCENE_CORNELL, 32, 32)
    _, _, st = ptref.OracleScene(sc.desc).render(sc.camera, sc.render_params(), 0, 2)
    assert int(st[3]) == 0
