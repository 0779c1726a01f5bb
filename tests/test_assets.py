"""Asset ingest of the host scene API (CPU): OBJ/MTL models with their
image textures and Radiance environments, read through pt_scene_build_obj,
against files written by tests/asset_files.py from the format
specifications."""
import numpy as np
import pytest

import ptamd
from asset_files import (bmp_bytes, jpeg_bytes, jpeg_reference_decode, png_bytes, ppm_bytes, rgbe_bytes,
                         rgbe_to_float, tga_bytes)

OBJ = """# two quads in the z = 0 plane, the second with negative indices
mtllib scene.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 2 0 0
v 2 1 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 1
usemtl textured
f 1/1/1 2/2/1 3/3/1 4/4/1
usemtl plain
f -5/-4/-1 -2/-3/-1 -1/-2/-1 -4/-1/-1
"""
MTL = """newmtl textured
Ka 0.1 0.1 0.1
map_Kd tex.png
newmtl plain
Ka 0.25 0.5 0.75
"""


def texels(desc, i):
    t = desc.textures[i]
    return np.ctypeslib.as_array(t.texels, shape=(t.height, t.width, 4)).copy()


def write_model(tmp_path, png):
    (tmp_path / "scene.obj").write_text(OBJ)
    (tmp_path / "scene.mtl").write_text(MTL)
    (tmp_path / "tex.png").write_bytes(png)


def test_obj_mtl_faces_materials_and_png_texture(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (5, 7, 3))
    write_model(tmp_path, png_bytes(img, 2, 8))
    sc = ptamd.Scene.from_obj("scene.obj", str(tmp_path) + "/", None, 64, 48)
    d = sc.desc
    assert (sc.info.width, sc.info.height) == (64, 48) and d.miss.type == ptamd.PT_MISS_GRADIENT
    assert d.num_meshes == 1
    m = d.meshes[0]
    # two quads fan-triangulated into four triangles of flattened corners
    assert m.num_triangles == 4 and m.num_vertices == 12
    v = np.ctypeslib.as_array(m.vertices, shape=(12, 3))
    uv = np.ctypeslib.as_array(m.uvs, shape=(12, 2))
    n = np.ctypeslib.as_array(m.normals, shape=(12, 3))
    np.testing.assert_array_equal(v[:6], [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_array_equal(v[6:], [[1, 0, 0], [2, 0, 0], [2, 1, 0], [1, 0, 0], [2, 1, 0], [1, 1, 0]])
    np.testing.assert_array_equal(uv[:3], [[0, 0], [1, 0], [1, 1]])
    np.testing.assert_array_equal(uv[6:9], [[0, 0], [1, 0], [1, 1]])
    assert np.all(n == [0, 0, 1])
    fm = np.ctypeslib.as_array(m.face_materials, shape=(4,))
    assert fm[0] == fm[1] and fm[2] == fm[3] and fm[0] != fm[2]
    tex_mat, plain_mat = d.materials[fm[0]], d.materials[fm[2]]
    assert tex_mat.type == plain_mat.type == ptamd.PT_MAT_LAMBERTIAN
    pt = d.textures[plain_mat.tex[0]]
    assert pt.type == ptamd.PT_TEX_CONSTANT and tuple(pt.rgb) == pytest.approx((0.25, 0.5, 0.75))
    t = d.textures[tex_mat.tex[0]]
    assert t.type == ptamd.PT_TEX_IMAGE and (t.width, t.height) == (7, 5)
    # byte * float32(1 / 255), rows flipped (row 0 = the image's bottom
    # row), alpha 255 for an RGB file
    rgba = np.concatenate([img[::-1], np.full((5, 7, 1), 255)], 2).astype(np.float32)
    np.testing.assert_array_equal(texels(d, tex_mat.tex[0]), rgba * np.float32(1.0 / 255.0))


@pytest.mark.parametrize("case", ["gray1", "gray2", "gray4", "gray8", "gray16", "graya8", "rgb16", "rgba8",
                                  "pal1", "pal4", "pal8"])
def test_png_colour_types_depths_and_filters(tmp_path, case):
    rng = np.random.default_rng(len(case))
    H, W = 6, 9
    palette = trns = None
    if case.startswith("gray") and case != "graya8":
        depth = int(case[4:])
        s = rng.integers(0, 1 << depth, (H, W, 1))
        top = s[..., 0] >> 8 if depth == 16 else (s[..., 0] * 255 // ((1 << depth) - 1))
        want = np.stack([top] * 3 + [np.full((H, W), 255)], 2)
        png = png_bytes(s, 0, depth)
    elif case == "graya8":
        s = rng.integers(0, 256, (H, W, 2))
        want = np.stack([s[..., 0]] * 3 + [s[..., 1]], 2)
        png = png_bytes(s, 4, 8)
    elif case == "rgb16":
        s = rng.integers(0, 65536, (H, W, 3))
        want = np.concatenate([s >> 8, np.full((H, W, 1), 255)], 2)
        png = png_bytes(s, 2, 16)
    elif case == "rgba8":
        s = rng.integers(0, 256, (H, W, 4))
        want = s
        png = png_bytes(s, 6, 8)
    else:
        depth = int(case[3:])
        npal = min(1 << depth, 11)
        palette = rng.integers(0, 256, (npal, 3))
        trns = [int(x) for x in rng.integers(0, 256, npal - 2)]
        s = rng.integers(0, npal, (H, W, 1))
        alpha = np.array(trns + [255, 255])[s[..., 0]]
        want = np.concatenate([palette[s[..., 0]], alpha[..., None]], 2)
        png = png_bytes(s, 3, depth, palette=palette, trns=trns)
    write_model(tmp_path, png)
    sc = ptamd.Scene.from_obj("scene.obj", str(tmp_path) + "/")
    d = sc.desc
    fm = np.ctypeslib.as_array(d.meshes[0].face_materials, shape=(4,))
    got = texels(d, d.materials[fm[0]].tex[0])
    expect = want[::-1].astype(np.float32) * np.float32(1.0 / 255.0)
    np.testing.assert_array_equal(got, expect)


def test_ppm_texture_and_bad_images(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (4, 3, 3)).astype(np.uint8)
    (tmp_path / "scene.obj").write_text(OBJ)
    (tmp_path / "scene.mtl").write_text(MTL.replace("tex.png", "tex.ppm"))
    (tmp_path / "tex.ppm").write_bytes(ppm_bytes(img))
    sc = ptamd.Scene.from_obj("scene.obj", str(tmp_path) + "/")  # owns the descriptor
    d = sc.desc
    fm = np.ctypeslib.as_array(d.meshes[0].face_materials, shape=(4,))
    got = texels(d, d.materials[fm[0]].tex[0])
    np.testing.assert_array_equal(got[..., :3], img[::-1].astype(np.float32) * np.float32(1.0 / 255.0))
    # truncated, unknown and missing images are refused with a message
    (tmp_path / "tex.ppm").write_bytes(ppm_bytes(img)[:-5])
    with pytest.raises(ptamd.PtError, match="truncated"):
        ptamd.Scene.from_obj("scene.obj", str(tmp_path) + "/")
    (tmp_path / "tex.ppm").write_bytes(b"GIF89a....")
    with pytest.raises(ptamd.PtError, match="images are read"):
        ptamd.Scene.from_obj("scene.obj", str(tmp_path) + "/")
    (tmp_path / "tex.ppm").unlink()
    with pytest.raises(ptamd.PtError, match="cannot read"):
        ptamd.Scene.from_obj("scene.obj", str(tmp_path) + "/")
    with pytest.raises(ptamd.PtError, match="obj"):
        ptamd.Scene.from_obj("missing.obj", str(tmp_path) + "/")


def test_radiance_hdr_environment(tmp_path):
    rng = np.random.default_rng(9)
    H, W = 6, 40
    rgbe = np.stack([rng.integers(0, 256, (H, W)) for _ in range(3)] + [rng.integers(100, 140, (H, W))], 2)
    rgbe[:, 5:20] = rgbe[:, 5:6]       # runs for the run-length encoder
    rgbe[2, :, 3] = 120                # a constant channel
    (tmp_path / "sky.hdr").write_bytes(rgbe_bytes(rgbe.astype(np.uint8)))
    write_model(tmp_path, png_bytes(np.zeros((2, 2, 3), int), 2, 8))
    sc = ptamd.Scene.from_obj("scene.obj", str(tmp_path) + "/", str(tmp_path / "sky.hdr"))
    d = sc.desc
    assert d.miss.type == ptamd.PT_MISS_HDR and d.miss.spherical == 1
    assert sc.info.flags & ptamd.PT_FLAG_ENV_MIS
    assert d.num_lights == 1 and d.lights[0].type == ptamd.PT_LIGHT_ENV
    got = texels(d, d.miss.texture)
    # rows in file order (the environment mapping's v = 0 is the file's top)
    np.testing.assert_array_equal(got[..., :3], rgbe_to_float(rgbe.astype(np.uint8)))


def texture_of(tmp_path, name, data):
    """load `data` as the map_Kd image `name` of the two-quad model"""
    (tmp_path / "scene.obj").write_text(OBJ)
    (tmp_path / "scene.mtl").write_text(MTL.replace("tex.png", name))
    (tmp_path / name).write_bytes(data)
    sc = ptamd.Scene.from_obj("scene.obj", str(tmp_path) + "/")
    d = sc.desc
    fm = np.ctypeslib.as_array(d.meshes[0].face_materials, shape=(4,))
    return texels(d, d.materials[fm[0]].tex[0])[::-1], sc  # back to file row order


@pytest.mark.parametrize("case", ["grey", "444", "420", "422", "420_restart", "separate_scans", "420_restart1"])
def test_jpeg_textures(tmp_path, case):
    """Baseline JPEG (ITU-T T.81) written by the test encoder: the decode
    equals a float64 reference decode of the same quantized coefficients
    to one level (rounding of different operation orders)."""
    rng = np.random.default_rng(sum(case.encode()))
    H, W = 21, 35
    yy, xx = np.mgrid[0:H, 0:W]
    smooth = np.stack([120 + 100 * np.sin(xx / 5.0 + k) * np.cos(yy / 7.0) for k in range(3)], 2)
    img = np.clip(smooth + rng.normal(0, 12, smooth.shape), 0, 255).astype(np.uint8)
    kw = dict(grey=dict(), **{"444": dict(sampling=(1, 1)), "420": dict(sampling=(2, 2)),
                             "422": dict(sampling=(2, 1)), "420_restart": dict(sampling=(2, 2), restart=2),
                             "420_restart1": dict(sampling=(2, 2), restart=1),
                             "separate_scans": dict(sampling=(2, 2), separate_scans=True, restart=3)})[case]
    src = img[..., 0] if case == "grey" else img
    data, comps, q, hv = jpeg_bytes(src, seed=7, **kw)
    want = jpeg_reference_decode(H, W, comps, q, hv)
    got, _ = texture_of(tmp_path, "tex.jpg", data)
    assert got.shape == (H, W, 4)
    diff = np.abs(np.round(got[..., :3] * 255).astype(int) - want.astype(int))
    assert diff.max() <= 1, diff.max()
    assert (diff == 0).mean() > 0.99
    # the decode is near the source (a lossy but faithful code)
    ref = np.repeat(src[..., None], 3, 2) if case == "grey" else src
    assert np.abs(want.astype(int) - ref.astype(int)).mean() < 12


def test_jpeg_refusals(tmp_path):
    img = np.full((8, 8, 3), 100, np.uint8)
    data, _, _, _ = jpeg_bytes(img)
    prog = data.replace(b"\xff\xc0", b"\xff\xc2", 1)
    with pytest.raises(ptamd.PtError, match="progressive"):
        texture_of(tmp_path, "tex.jpg", prog)
    with pytest.raises(ptamd.PtError, match="JPEG"):
        texture_of(tmp_path, "tex.jpg", data[: len(data) // 2])


@pytest.mark.parametrize("fmt", ["tga_rgb", "tga_rgba_rle_top", "tga_grey_rle", "bmp24", "bmp32_top"])
def test_tga_bmp_textures(tmp_path, fmt):
    rng = np.random.default_rng(len(fmt))
    H, W = 6, 5
    ch = 4 if "rgba" in fmt or "32" in fmt else 3
    img = rng.integers(0, 256, (H, W, ch)).astype(np.uint8)
    img[2, 1:4] = img[2, 1]  # a run for the RLE coder
    if fmt == "tga_grey_rle":
        img = img[..., 0]
        data, name, want = tga_bytes(img, rle=True), "tex.tga", np.repeat(img[..., None], 3, 2)
    elif fmt.startswith("tga"):
        data, name, want = tga_bytes(img, rle="rle" in fmt, top_down="top" in fmt), "tex.tga", img
    else:
        data, name, want = bmp_bytes(img, top_down="top" in fmt), "tex.bmp", img
    got, _ = texture_of(tmp_path, name, data)
    np.testing.assert_array_equal(got[..., : want.shape[2]], want.astype(np.float32) * np.float32(1.0 / 255.0))
    if want.shape[2] == 3:
        assert (got[..., 3] == 1.0).all()


def test_malformed_images_are_refused_or_read(tmp_path):
    """Random byte mutations and truncations of valid JPEG, TGA and BMP files
    either decode or raise PtError; the reader never reads out of bounds
    (a fault would end the test process)."""
    rng = np.random.default_rng(40)
    img = rng.integers(0, 256, (13, 17, 3)).astype(np.uint8)
    files = {"tex.jpg": jpeg_bytes(img, sampling=(2, 2), restart=2)[0], "tex.tga": tga_bytes(img, rle=True),
             "tex.bmp": bmp_bytes(img)}
    outcomes = {"ok": 0, "refused": 0}
    for name, good in files.items():
        for trial in range(60):
            data = bytearray(good)
            if trial % 3 == 0:
                data = data[: int(rng.integers(1, len(data)))]
            else:
                for _ in range(int(rng.integers(1, 6))):
                    data[int(rng.integers(0, len(data)))] = int(rng.integers(0, 256))
            try:
                texture_of(tmp_path, name, bytes(data))
                outcomes["ok"] += 1
            except ptamd.PtError:
                outcomes["refused"] += 1
    assert outcomes["refused"] > 20 and sum(outcomes.values()) == 180
