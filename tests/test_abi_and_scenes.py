"""Host-side checks that need no GPU: the C ABI of libpt_hip.so (exports,
struct layouts, version, error path without a device) and the scene
generators of the host scene API (pt_scene.h), which restate the
reference's RTIOW final scene (host_includes/scenes.hpp) and Cornell box."""
import ctypes as C
import math
import re
import subprocess

import numpy as np
import pytest

import ptamd

REPO = ptamd.REPO


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(ptamd.LIB_PATH))
    names = ptamd.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    nm = subprocess.run(["nm", "-D", "--defined-only", str(ptamd.LIB_PATH)], capture_output=True, text=True,
                        check=True).stdout
    exported = set(re.findall(r"\bT (pt_\w+)", nm))
    assert set(names) <= exported


def test_api_version_matches_header():
    text = (REPO / "include" / "pt_api.h").read_text()
    want = int(re.search(r"#define PT_API_VERSION (\d+)", text).group(1))
    assert ptamd.load().pt_api_version() == want


STRUCTS = {
    "pt_sphere_desc": ptamd.SphereDesc, "pt_rect_desc": ptamd.RectDesc, "pt_box_desc": ptamd.BoxDesc,
    "pt_mesh_desc": ptamd.MeshDesc, "pt_instance_desc": ptamd.InstanceDesc,
    "pt_material_desc": ptamd.MaterialDesc, "pt_texture_desc": ptamd.TextureDesc,
    "pt_light_desc": ptamd.LightDesc, "pt_miss_desc": ptamd.MissDesc, "pt_scene_desc": ptamd.SceneDesc,
    "pt_camera": ptamd.Camera, "pt_render_params": ptamd.RenderParams, "pt_render_stats": ptamd.RenderStats,
    "pt_accel_stats": ptamd.AccelStats, "pt_scene_info": ptamd.SceneInfo, "pt_tune_result": ptamd.TuneResult,
}


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof and every field offset as gcc lays out include/*.h, against
    the ctypes mirror."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pt_api.h"', '#include "pt_scene.h"',
             'int main(void) {']
    for cname, py in STRUCTS.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['  printf("pt_ray sizeof %zu\\n", sizeof(pt_ray));', '  printf("pt_hit sizeof %zu\\n", sizeof(pt_hit));',
              '  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", str(REPO / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            a, b, c = ln.split()
            got[(a, b)] = int(c)
    for cname, py in STRUCTS.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
    assert got[("pt_ray", "sizeof")] == ptamd.RAY_DTYPE.itemsize == 32
    assert got[("pt_hit", "sizeof")] == ptamd.HIT_DTYPE.itemsize == 16


def test_create_reports_missing_device():
    """Without a usable GPU pt_create fails with a status and a message; it
    never falls back to a CPU path."""
    try:
        ctx = ptamd.Context(0)
    except ptamd.PtError as e:
        assert "device" in str(e).lower()
    else:  # a GPU is present: nothing to check here
        ctx.close()


def test_invalid_arguments_are_rejected():
    lib = ptamd.load()
    assert lib.pt_set_scene(None, None) == -1
    assert lib.pt_render(None, None, 0, 0, None) == -1
    assert lib.pt_trace_closest(None, None, 0, None) == -1
    h = C.c_void_p()
    assert lib.pt_scene_build(99, 0, 0, 0, C.byref(h)) != 0


# -------------------------------------------------------- scene builders
def o2w(inst):
    return np.array(inst.obj_to_world[:], np.float64).reshape(3, 4)


def w2o(inst):
    return np.array(inst.world_to_obj[:], np.float64).reshape(3, 4)


def rot_y(deg):
    a = math.radians(deg)
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


def test_cornell_box_scene():
    s = ptamd.Scene(ptamd.PT_SCENE_CORNELL)
    d, i, cam = s.desc, s.info, s.camera
    assert (i.width, i.height, i.spp, i.max_depth, i.russian, i.rr_min_depth) == (512, 512, 64, 16, 1, 3)
    assert (d.num_rects, d.num_boxes, d.num_instances, d.num_lights) == (6, 2, 3, 1)
    L = d.lights[0]
    assert (L.type, L.axis, L.a0, L.a1, L.b0, L.b1, L.k) == (ptamd.PT_LIGHT_RECT, ptamd.PT_AXIS_Y, 213, 343, 227,
                                                             332, 554)
    assert list(L.emission) == [15, 15, 15]
    # boxes: rotate about Y, then translate (transforms.hpp:51-419)
    want = [(rot_y(15), (265, 0, 295)), (rot_y(-18), (130, 0, 65))]
    boxes = [d.instances[k] for k in range(3) if d.instances[k].kind == ptamd.PT_GEOM_BOXES]
    for inst, (R, t) in zip(boxes, want):
        M = o2w(inst)
        np.testing.assert_allclose(M[:, :3], R, atol=1e-6)
        np.testing.assert_allclose(M[:, 3], t, atol=1e-4)
        full = np.vstack([M, [0, 0, 0, 1]])
        inv = np.vstack([w2o(inst), [0, 0, 0, 1]])
        np.testing.assert_allclose(inv @ full, np.eye(4), atol=1e-5)
    sizes = sorted(tuple(np.array(d.boxes[k].pmax[:]) - np.array(d.boxes[k].pmin[:])) for k in range(2))
    assert sizes == [(165, 165, 165), (165, 330, 165)]
    # camera (278, 278, -800) looking at (278, 278, 0), vfov 40; image left is +x
    assert list(cam.origin) == [278, 278, -800] and cam.lens_radius == 0
    assert cam.horizontal[0] < 0
    # camera.hpp:13-51 with focus distance 10: vertical = 2 tan(vfov/2) * 10
    assert abs(cam.vertical[1] - 20 * math.tan(math.radians(20))) < 1e-4
    # the green wall (x = 555) and the red wall (x = 0)
    walls = {(d.rects[k].axis, d.rects[k].k): d.rects[k] for k in range(6)}
    green = d.textures[d.materials[walls[(0, 555.0)].material].tex[0]].rgb[:]
    red = d.textures[d.materials[walls[(0, 0.0)].material].tex[0]].rgb[:]
    assert green[1] > green[0] and red[0] > red[1]


def test_random_spheres_scene():
    s = ptamd.Scene(ptamd.PT_SCENE_RANDOM_SPHERES)
    d, i = s.desc, s.info
    assert (i.width, i.height, i.spp, i.max_depth, i.russian) == (400, 267, 16, 50, 1)
    n = d.num_spheres
    sph = [(np.array(d.spheres[k].center[:]), d.spheres[k].radius, d.spheres[k].material) for k in range(n)]
    assert sph[0][1] == 1000 and list(sph[0][0]) == [0, -1000, 0]
    big = {tuple(c): r for c, r, _ in sph[1:4]} if n > 3 else {}
    small = [x for x in sph if x[1] == np.float32(0.2)]
    assert len(small) + 4 == n and 300 < len(small) <= 484
    for c, r, m in small:
        assert c[1] == np.float32(0.2)
        assert np.linalg.norm(c - np.array([4, 0.2, 0])) > 0.9
    types = {d.materials[m].type for _, _, m in sph}
    assert types == {ptamd.PT_MAT_LAMBERTIAN, ptamd.PT_MAT_METAL, ptamd.PT_MAT_DIELECTRIC}
    # deterministic: the scene RNG is seeded
    s2 = ptamd.Scene(ptamd.PT_SCENE_RANDOM_SPHERES)
    assert all(s2.desc.spheres[k].center[:] == d.spheres[k].center[:] for k in range(n))
    assert len(big) == 3


def test_batch_workload_scenes():
    """The 64-scene batch: scenes 0..31 hold 50..5000 spheres (plus the
    ground), 32..63 hold 10k..2M triangles under an environment light;
    square resolutions of side 256..1024; 4, 16 or 64 spp; deterministic."""
    sides, spps = [], set()
    for i in (0, 1, 17, 31, 32, 45, 63):
        s = ptamd.Scene(ptamd.PT_SCENE_BATCH, 0, 0, i)
        d, info = s.desc, s.info
        assert info.width == info.height and 256 <= info.width <= 1024
        assert info.spp in (4, 16, 64)
        sides.append(info.width)
        spps.add(info.spp)
        if i < 32:
            assert d.num_meshes == 0 and 51 <= d.num_spheres <= 5001
            assert d.miss.type == ptamd.PT_MISS_GRADIENT and d.num_lights == 0
        else:
            tris = sum(d.meshes[k].num_triangles for k in range(d.num_meshes))
            assert d.num_spheres == 0 and 5000 <= tris <= 2_500_000
            assert d.num_lights == 1 and d.lights[0].type == ptamd.PT_LIGHT_ENV
            assert info.flags & ptamd.PT_FLAG_ENV_MIS
        again = ptamd.Scene(ptamd.PT_SCENE_BATCH, 0, 0, i)
        assert again.info.width == info.width and again.desc.num_spheres == d.num_spheres
        if d.num_spheres:
            k = d.num_spheres - 1
            assert again.desc.spheres[k].center[:] == d.spheres[k].center[:]
    assert len(set(sides)) > 3 and len(spps) > 1
    with pytest.raises(ptamd.PtError):
        ptamd.Scene(ptamd.PT_SCENE_BATCH, 0, 0, 64)
