"""The reference's own scene file, unchanged, on this package's host API.

host_includes/scenes.hpp is compiled as it lies in the reference tree (a
symlink in a temporary folder points at it, so its #include "camera.hpp"
etc. resolve to optix-path-tracer_amd/host/) together with
tests/ref_scenes_driver.cpp, linked against libpt_hip.so, and run: each scene
function records its scene graph through the facade, which flattens it as the
first Context::launch would.  The flattened descriptors are checked against
what the scene functions state (scenes.hpp:17-93 InOneWeekend, :188-260
Cornell) and against the thin-lens camera of
camera.hpp:10-38 evaluated in float64.  No GPU is touched; skipped when the
reference tree or the built library is absent.
"""
import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

import ptamd

REPO = Path(__file__).resolve().parent.parent
REF_SCENES = Path("/root/reference/OptiX-Path-Tracer/host_includes/scenes.hpp")
LIB = REPO / "optix-path-tracer_amd" / "libpt_hip.so"

pytestmark = pytest.mark.skipif(not (REF_SCENES.exists() and LIB.exists()),
                                reason="reference tree or libpt_hip.so absent")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("refscenes")
    (d / "scenes.hpp").symlink_to(REF_SCENES)
    exe = d / "ref_scenes_driver"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(REPO / "optix-path-tracer_amd" / "host"),
                    "-I", str(REPO / "include"), "-I", str(d), "-o", str(exe),
                    str(REPO / "tests" / "ref_scenes_driver.cpp"), "-L", str(LIB.parent), "-lpt_hip", "-lz",
                    f"-Wl,-rpath,{LIB.parent}"], check=True)

    def run(scene, W=500, H=500):
        out = subprocess.run([str(exe), scene, str(W), str(H)], check=True, capture_output=True, text=True,
                             timeout=60)
        return json.loads(out.stdout.strip().splitlines()[-1])

    return run


def thin_lens(lookfrom, lookat, up, vfov, aspect, aperture, dist):
    """camera.hpp:10-38 (Shirley's camera) in float64."""
    o, at, up = (np.array(v, np.float64) for v in (lookfrom, lookat, up))
    hh = math.tan(math.radians(vfov) / 2)
    hw = aspect * hh
    w = (o - at) / np.linalg.norm(o - at)
    u = np.cross(up, w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    return {"origin": o, "lower_left": o - hw * dist * u - hh * dist * v - dist * w,
            "horizontal": 2 * hw * dist * u, "vertical": 2 * hh * dist * v, "lens_radius": aperture / 2}


def check_camera(cam, want):
    for k in ("origin", "lower_left", "horizontal", "vertical"):
        np.testing.assert_allclose(cam[k], want[k], rtol=1e-5, atol=1e-4, err_msg=k)
    assert math.isclose(cam["lens_radius"], want["lens_radius"], rel_tol=1e-6)


def tex_rgb(d, mat):
    return d["textures"][d["materials"][mat]["tex"][0]]["rgb"]


def test_cornell_flattens_as_the_scene_states(driver):
    d = driver("Cornell", 400, 300)
    # six walls and the light (scenes.hpp:221-227), the green box (:231-234)
    want_rects = [(ptamd.PT_AXIS_X, 1, [0, 555, 0, 555, 555]), (ptamd.PT_AXIS_X, 0, [0, 555, 0, 555, 0]),
                  (ptamd.PT_AXIS_Y, 1, [213, 343, 227, 332, 554]), (ptamd.PT_AXIS_Y, 1, [0, 555, 0, 555, 555]),
                  (ptamd.PT_AXIS_Y, 0, [0, 555, 0, 555, 0]), (ptamd.PT_AXIS_Z, 1, [0, 555, 0, 555, 555])]
    assert [(r["axis"], r["flip"], r["range"]) for r in d["rects"]] == want_rects
    red, green, white = (0.65, 0.05, 0.05), (0.12, 0.45, 0.15), (0.73, 0.73, 0.73)
    want_cols = [red, green, (7, 7, 7), white, white, white]
    for r, c in zip(d["rects"], want_cols):
        np.testing.assert_allclose(tex_rgb(d, r["material"]), c, rtol=1e-6)
    light_mat = d["materials"][d["rects"][2]["material"]]
    assert light_mat["type"] == ptamd.PT_MAT_DIFFUSE_LIGHT
    assert {d["materials"][r["material"]]["type"] for i, r in enumerate(d["rects"]) if i != 2} == {
        ptamd.PT_MAT_LAMBERTIAN}
    (box,) = d["boxes"]
    assert box["pmin"] == [0, 0, 0] and box["pmax"] == [165, 330, 165]
    np.testing.assert_allclose(tex_rgb(d, box["material"]), green, rtol=1e-6)
    # the box's instance: rotated 15 degrees about Y, then placed at (265, 0, 295)
    inst = [i for i in d["instances"] if i["kind"] == ptamd.PT_GEOM_BOXES]
    assert len(inst) == 1 and inst[0]["has_transform"]
    m = np.array(inst[0]["obj_to_world"]).reshape(3, 4)
    c, s = math.cos(math.radians(15)), math.sin(math.radians(15))
    np.testing.assert_allclose(m[:, :3], [[c, 0, s], [0, 1, 0], [-s, 0, c]], atol=1e-6)
    np.testing.assert_allclose(m[:, 3], [265, 0, 295], atol=1e-4)
    # Rectangle_PDF light with emission 7 (:192-196), dark constant miss (:200)
    (lt,) = d["lights"]
    assert lt["type"] == ptamd.PT_LIGHT_RECT and lt["axis"] == ptamd.PT_AXIS_Y
    assert lt["range"] == [213, 343, 227, 332, 554] and lt["emission"] == [7, 7, 7]
    assert d["miss"]["type"] == ptamd.PT_MISS_CONSTANT
    check_camera(d["camera"], thin_lens((278, 278, -800), (278, 278, 0), (0, 1, 0), 40, 400 / 300, 0, 10))


def test_in_one_weekend_flattens_as_the_scene_states(driver):
    d = driver("InOneWeekend", 400, 267)
    sp = d["spheres"]
    # ground, 22 x 22 small spheres, three large ones (scenes.hpp:39-72)
    assert len(sp) == 1 + 22 * 22 + 3
    assert sp[0]["center"] == [0, -1000, -1] and sp[0]["radius"] == 1000
    np.testing.assert_allclose(tex_rgb(d, sp[0]["material"]), (0.5, 0.5, 0.5))
    kinds = {ptamd.PT_MAT_LAMBERTIAN: 0, ptamd.PT_MAT_METAL: 0, ptamd.PT_MAT_DIELECTRIC: 0}
    k = 1
    for a in range(-11, 11):
        for b in range(-11, 11):
            s = sp[k]
            k += 1
            x, y, z = s["center"]
            assert a <= x < a + 1 and y == np.float32(0.2) and b <= z < b + 1 and s["radius"] == np.float32(0.2)
            mat = d["materials"][s["material"]]
            kinds[mat["type"]] += 1
            rgb = np.array(tex_rgb(d, s["material"]))
            if mat["type"] == ptamd.PT_MAT_LAMBERTIAN:
                assert np.all((rgb >= 0) & (rgb < 1))
            elif mat["type"] == ptamd.PT_MAT_METAL:
                assert np.all((rgb >= 0.5) & (rgb < 1)) and 0 <= mat["param"][0] < 0.5
    # choose_mat thresholds 0.8 / 0.95 over 484 draws (binomial, 4.5 sigma)
    n = 484
    for t, p in ((ptamd.PT_MAT_LAMBERTIAN, 0.8), (ptamd.PT_MAT_METAL, 0.15), (ptamd.PT_MAT_DIELECTRIC, 0.05)):
        assert abs(kinds[t] - n * p) <= 4.5 * math.sqrt(n * p * (1 - p)), kinds
    big = sp[-3:]
    assert [s["center"] for s in big] == [[4, 1, 0], [0, 1, 0.5], [-4, 1, 1]] and all(s["radius"] == 1 for s in big)
    types = [d["materials"][s["material"]]["type"] for s in big]
    assert types == [ptamd.PT_MAT_DIELECTRIC, ptamd.PT_MAT_LAMBERTIAN, ptamd.PT_MAT_METAL]
    np.testing.assert_allclose(tex_rgb(d, big[1]["material"]), (0.4, 0.2, 0.1), rtol=1e-6)
    np.testing.assert_allclose(tex_rgb(d, big[2]["material"]), (0.7, 0.6, 0.5), rtol=1e-6)
    assert d["materials"][big[2]["material"]]["param"][0] == 0
    assert d["lights"] == []
    assert d["miss"]["type"] == ptamd.PT_MISS_GRADIENT
    np.testing.assert_allclose(d["miss"]["color1"], (1, 1, 1))
    np.testing.assert_allclose(d["miss"]["color2"], (0.5, 0.7, 1), rtol=1e-6)
    check_camera(d["camera"], thin_lens((13, 2, 3), (0, 0, 0), (0, 1, 0), 20, 400 / 267, 0.1, 10))


def test_final_next_week_refuses_its_missing_earth_map(driver):
    """Final_Next_Week's Earth sphere reads assets/other_textures/map.jpg
    (scenes.hpp:332-335), which the reference tree does not hold: the
    facade's Image_Texture refuses the scene loudly instead of rendering it
    without the texture."""
    with pytest.raises(subprocess.CalledProcessError) as e:
        driver("Final_Next_Week", 300, 300)
    assert "map.jpg" in e.value.stderr
