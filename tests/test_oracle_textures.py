"""Pins the oracle's procedural textures with closed forms, and the host
texture classes' tables with the reference's construction order.

Reference code followed: checkered_texture.cu:8-15, noise_texture.cu:12-84
(Perlin noise, 7-octave turbulence; the axis switch falls through to the z
term), gradient_texture.cu:7-14, host_includes/textures.hpp:39-112 (Perlin
tables drawn from rnd(): 256 unit vectors, then perm_x, perm_y, perm_z by a
Fisher-Yates pass from i = 255 down), scenes.hpp:95-186 (MovingSpheres).

Noise closed forms: at a lattice point every corner weight is 0 or has a
zero offset vector, so the noise is 0; at (i + 1/4, j, k) only the two x
corners contribute, (1 - s) c0.x / 4 - 3 s c1.x / 4 with s = 5/32 the
smoothstep of 1/4, and at (2i + 1/2, 2j, 2k) (c0.x - c1.x) / 4; octaves from
4p on sit on the lattice.  These fix the hash (which permutation indexes
which axis), the offset signs, the smoothstep and the octave weights.
"""
import math

import numpy as np
import pytest

import ptamd
import ptref
from helpers import PyScene


def noise_tables(seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(256, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    perms = np.stack([rng.permutation(256) for _ in range(3)])
    return v.astype(np.float32), perms


def corner(ranvec, perms, i, j, k):
    return ranvec[perms[0][i & 255] ^ perms[1][j & 255] ^ perms[2][k & 255]].astype(np.float64)


def quarter_noise(ranvec, perms, lat, axis):
    """noise at lat + e_axis / 4 (lat integer)"""
    s = 0.25 * 0.25 * (3 - 0.5)
    hi = list(lat)
    hi[axis] += 1
    return (1 - s) * 0.25 * corner(ranvec, perms, *lat)[axis] - s * 0.75 * corner(ranvec, perms, *hi)[axis]


def tex_value(z, turb):
    """0.5 (1 + sin(z + 5 turb)) with the sine's argument rounded to float32
    as the texture forms it (|z| up to 600: an ulp there is 6e-5)"""
    arg = np.float32(z) + np.float32(5) * np.float32(turb)
    return 0.5 * (1 + math.sin(float(arg)))


def half_noise(ranvec, perms, lat, axis):
    hi = list(lat)
    hi[axis] += 1
    return 0.25 * (corner(ranvec, perms, *lat)[axis] - corner(ranvec, perms, *hi)[axis])


@pytest.fixture(scope="module")
def tex_scene():
    s = PyScene()
    ranvec, perms = noise_tables(5)
    a = s.texture((0.1, 0.2, 0.3))
    b = s.texture((0.7, 0.8, 0.9))
    img = np.random.default_rng(6).uniform(0, 1, (4, 8, 4)).astype(np.float32)
    im = s.texture(image=img)
    ids = {
        "a": a, "b": b, "img": im,
        "checker": s.checker(a, b),
        "checker_img": s.checker(im, b),
        "noise1": s.noise(1.0, ranvec, perms),
        "noise4": s.noise(4.0, ranvec, perms),
        "gradient": s.gradient((1.0, 0.0, 0.0), (0.0, 0.5, 0.0), (0.0, 0.0, 2.0)),
    }
    s.spheres.append(((0, 0, 0), 1.0, s.material(ptamd.PT_MAT_LAMBERTIAN, a)))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    orc = ptref.OracleScene(s.desc())
    yield orc, ids, ranvec, perms
    orc.close()


def test_checker_picks_by_sign_of_sines(tex_scene):
    orc, ids, _, _ = tex_scene
    rng = np.random.default_rng(1)
    A, B = np.float32([0.1, 0.2, 0.3]), np.float32([0.7, 0.8, 0.9])
    seen = set()
    for P in rng.uniform(-3, 3, (400, 3)).astype(np.float32):
        s = math.sin(10 * P[0]) * math.sin(10 * P[1]) * math.sin(10 * P[2])
        if abs(s) < 1e-4:
            continue
        got = orc.texture(ids["checker"], 0.3, 0.6, P)
        np.testing.assert_array_equal(got, A if s < 0 else B)
        seen.add(s < 0)
    assert seen == {True, False}


def test_checker_passes_uv_to_its_child(tex_scene):
    orc, ids, _, _ = tex_scene
    rng = np.random.default_rng(2)
    n = 0
    for P in rng.uniform(-3, 3, (200, 3)).astype(np.float32):
        s = math.sin(10 * P[0]) * math.sin(10 * P[1]) * math.sin(10 * P[2])
        if s > -1e-4:
            continue
        u, v = rng.uniform(0, 1, 2)
        np.testing.assert_array_equal(orc.texture(ids["checker_img"], u, v, P), orc.texture(ids["img"], u, v, P))
        n += 1
    assert n > 50
    # the image itself varies with (u, v)
    assert not np.array_equal(orc.texture(ids["img"], 0.1, 0.1, [0, 0, 0]), orc.texture(ids["img"], 0.6, 0.4, [0, 0, 0]))


def test_gradient_weights_by_abs_unit_point(tex_scene):
    orc, ids, _, _ = tex_scene
    rng = np.random.default_rng(3)
    C = np.array([[1.0, 0, 0], [0, 0.5, 0], [0, 0, 2.0]])
    for P in rng.normal(size=(300, 3)).astype(np.float32) * 7:
        n = np.abs(P.astype(np.float64) / np.linalg.norm(P.astype(np.float64)))
        np.testing.assert_allclose(orc.texture(ids["gradient"], 0, 0, P), n @ C, rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(orc.texture(ids["gradient"], 0, 0, [0, -3, 0]), [0, 0.5, 0], atol=1e-7)


def test_noise_zero_on_lattice(tex_scene):
    """turb = 0 on integer points: the texture is 0.5 (1 + sin(scale z))"""
    orc, ids, _, _ = tex_scene
    rng = np.random.default_rng(4)
    for P in rng.integers(-300, 300, (200, 3)).astype(np.float32):
        want = 0.5 * (1 + math.sin(P[2]))
        np.testing.assert_allclose(orc.texture(ids["noise1"], 0, 0, P), [want] * 3, atol=2e-6)
        Q = P / np.float32(4)  # scale 4 maps it back onto the lattice
        np.testing.assert_allclose(orc.texture(ids["noise4"], 0, 0, Q), [want] * 3, atol=2e-6)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_noise_quarter_points(tex_scene, axis):
    orc, ids, ranvec, perms = tex_scene
    rng = np.random.default_rng(10 + axis)
    for lat in rng.integers(-300, 300, (120, 3)):
        P = lat.astype(np.float64)
        P[axis] += 0.25
        n1 = quarter_noise(ranvec, perms, lat, axis)
        n2 = half_noise(ranvec, perms, 2 * lat, axis)
        want = tex_value(P[2], abs(n1 + 0.5 * n2))
        np.testing.assert_allclose(orc.texture(ids["noise1"], 0, 0, P.astype(np.float32)), [want] * 3, atol=3e-6)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_noise_half_points(tex_scene, axis):
    orc, ids, ranvec, perms = tex_scene
    rng = np.random.default_rng(20 + axis)
    for lat in rng.integers(-300, 300, (120, 3)):
        P = lat.astype(np.float64) * 2
        P[axis] += 0.5
        want = tex_value(P[2], abs(half_noise(ranvec, perms, 2 * lat, axis)))
        np.testing.assert_allclose(orc.texture(ids["noise1"], 0, 0, P.astype(np.float32)), [want] * 3, atol=3e-6)
        # the same value at scale 4 from a quarter of the point
        np.testing.assert_allclose(orc.texture(ids["noise4"], 0, 0, (P / 4).astype(np.float32)), [want] * 3,
                                   atol=3e-6)


# ------------------------------------------------------ host texture classes
class HostRnd:
    """rnd() of the host scene code: std::mt19937 seeded with `seed`, each
    draw the top 24 bits of one output times 2^-24."""

    def __init__(self, seed):
        self.g = np.random.MT19937()
        self.g._legacy_seeding(seed)

    def __call__(self):
        return np.float32(int(self.g.random_raw()) >> 8) * np.float32(2.0 ** -24)


def moving_spheres_noise(seed):
    r = HostRnd(seed)
    third, two_thirds = np.float32(1) / np.float32(3), np.float32(2) / np.float32(3)
    for _ in range(22 * 22):
        cm = r()
        r(), r(), r()  # centre x, z and the unused centre2 offset
        r(), r(), r()  # tint
        if third <= cm < two_thirds:
            r()  # metal fuzz
    r(), r(), r()  # glass tint
    vec = np.zeros((256, 3), np.float32)
    for i in range(256):
        x, y, z = (np.float32(-1) + np.float32(2) * r() for _ in range(3))
        l = np.sqrt(x * x + y * y + z * z)
        vec[i] = (x / l, y / l, z / l)
    perms = []
    for _ in range(3):
        p = list(range(256))
        for i in range(255, 0, -1):
            t = int(r() * np.float32(i + 1))
            p[i], p[t] = p[t], p[i]
        perms.append(p)
    return vec, np.array(perms)


def textures_of(desc, kind):
    return [desc.textures[i] for i in range(desc.num_textures) if desc.textures[i].type == kind]


@pytest.mark.parametrize("seed", [0, 7])
def test_moving_spheres_textures(seed):
    sc = ptamd.Scene(ptamd.PT_SCENE_MOVING_SPHERES, 40, 27, seed)
    d = sc.desc
    (chk,) = textures_of(d, ptamd.PT_TEX_CHECKER)
    np.testing.assert_allclose(d.textures[chk.width].rgb[:], [0.2, 0.3, 0.1])
    np.testing.assert_allclose(d.textures[chk.height].rgb[:], [0.9, 0.9, 0.9])
    (grad,) = textures_of(d, ptamd.PT_TEX_GRADIENT)
    assert np.ctypeslib.as_array(grad.texels, (12,)).reshape(3, 4)[:, 3].tolist() == [0, 0, 0]
    (noise,) = textures_of(d, ptamd.PT_TEX_NOISE)
    assert noise.rgb[0] == 4.0
    t = np.ctypeslib.as_array(noise.texels, (ptamd.PT_NOISE_FLOATS,))
    vec = t[:1024].reshape(256, 4)
    np.testing.assert_allclose(np.linalg.norm(vec[:, :3].astype(np.float64), axis=1), 1, atol=1e-6)
    assert (vec[:, 3] == 0).all()
    perms = t[1024:].reshape(3, 256)
    for p in perms:
        assert sorted(p.astype(int).tolist()) == list(range(256))
    want_vec, want_perms = moving_spheres_noise(seed)
    np.testing.assert_array_equal(vec[:, :3], want_vec)
    np.testing.assert_array_equal(perms, want_perms)
    assert d.num_spheres == 1 + 22 * 22 + 3 and d.num_rects == 1
    assert sc.info.max_depth == 50


def test_moving_spheres_oracle_render():
    sc = ptamd.Scene(ptamd.PT_SCENE_MOVING_SPHERES, 24, 16)
    orc = ptref.OracleScene(sc.desc)
    acc, _, _ = orc.render(sc.camera, sc.render_params(), 0, 1)
    acc = np.asarray(acc)
    assert np.isfinite(acc).all() and acc.max() > 0
    orc.close()


def test_vector_texture_selects_by_face_index(tex_scene):
    """vector.cu:7-12: child i of the list, black out of range"""
    s = PyScene()
    a = s.texture((0.1, 0.2, 0.3))
    b = s.texture((0.7, 0.8, 0.9))
    img = np.random.default_rng(9).uniform(0, 1, (4, 8, 4)).astype(np.float32)
    im = s.texture(image=img)
    chk = s.checker(a, b)
    data = np.float32([b, im, chk, a])
    s.textures.append((ptamd.PT_TEX_VECTOR, (0, 0, 0), None, (4, 0), data))
    vec = len(s.textures) - 1
    s.spheres.append(((0, 0, 0), 1.0, s.material(ptamd.PT_MAT_LAMBERTIAN, vec)))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    orc = ptref.OracleScene(s.desc())
    P = np.float32([0.31, -0.72, 1.17])
    for i, child in enumerate([b, im, chk, a]):
        np.testing.assert_array_equal(orc.texture(vec, 0.3, 0.6, P, i), orc.texture(child, 0.3, 0.6, P))
    for i in (-1, 4, 100):
        np.testing.assert_array_equal(orc.texture(vec, 0.3, 0.6, P, i), np.zeros(3, np.float32))
    orc.close()
