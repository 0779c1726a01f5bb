"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

  * acceleration structure: Morton codes, sorted order and every 128-byte
    BVH4 node bit-exact against the oracle's sequential LBVH build;
  * closest-hit queries: instance and primitive ids bit-exact, hit distance
    equal (tolerance 1e-5 relative written below), node/primitive fetch
    counts equal, against the oracle's BVH traversal and brute force;
  * occlusion queries under the shared shadow-ray rule;
  * per-path radiance: paths whose discrete decisions match the oracle's
    (signature) agree within 1e-4 relative; accumulated images agree.
"""
import sys

import numpy as np
import pytest

import ptamd
import ptref
from helpers import PyScene, camera_rays, icosphere, random_rays, same_hits

pytestmark = pytest.mark.gpu

T_REL = 1e-5      # hit distance tolerance (relative)
RAD_REL = 1e-4    # per-path radiance tolerance for non-divergent paths


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ptamd.Context(0)
    yield c
    c.close()


def gpu_scene(ctx, desc):
    ctx.set_scene(desc)
    return ctx.build_accel()


def sphere_soup(n, seed):
    rng = np.random.default_rng(seed)
    s = PyScene()
    m = s.lambert((0.5, 0.5, 0.5))
    for _ in range(n):
        c = rng.uniform(-10, 10, 3)
        s.spheres.append((tuple(c), float(rng.uniform(0.05, 0.8)), m))
    s.instance(ptamd.PT_GEOM_SPHERES, 0, n)
    return s


def mesh_scene(subdiv, copies, seed):
    rng = np.random.default_rng(seed)
    s = PyScene()
    m = s.lambert((0.7, 0.7, 0.7))
    v, f = icosphere(subdiv)
    s.meshes.append((v, f, None, None, m))
    for i in range(copies):
        sc = rng.uniform(0.5, 2.0)
        ang = rng.uniform(0, 2 * np.pi)
        c, sn = np.cos(ang), np.sin(ang)
        R = np.array([[c, 0, sn], [0, 1, 0], [-sn, 0, c]]) * sc
        t = rng.uniform(-10, 10, 3)
        s.instance(ptamd.PT_GEOM_MESH, 0, 0, np.hstack([R, t[:, None]]))
    return s


def scene_desc(name, keep):
    # the descriptors point into their Scene / builder, kept alive in `keep`
    if name in ("random_spheres", "cornell"):
        sc = ptamd.Scene(ptamd.PT_SCENE_RANDOM_SPHERES if name == "random_spheres" else ptamd.PT_SCENE_CORNELL)
        keep.append(sc)
        return sc.desc
    if name == "sphere_soup_3000":
        s = sphere_soup(3000, 7)
    elif name == "sphere_soup_5":
        s = sphere_soup(5, 8)
    elif name == "icosphere_x6":
        s = mesh_scene(4, 6, 3)
    else:
        raise KeyError(name)
    keep.append(s)
    return s.desc()


ACCEL_CASES = ["random_spheres", "cornell", "sphere_soup_3000", "sphere_soup_5", "icosphere_x6"]


@pytest.mark.parametrize("name", ACCEL_CASES)
def test_accel_bit_exact(ctx, name):
    keep = []
    desc = scene_desc(name, keep)
    say(f"[gpu] accel build {name}")
    gpu_scene(ctx, desc)
    g_nodes, g_order, g_morton, g_info = ctx.read_accel()
    o_nodes, o_order, o_morton, o_info = ptref.OracleScene(desc).accel()
    assert np.array_equal(g_info, o_info)
    assert np.array_equal(g_morton, o_morton)
    assert np.array_equal(g_order, o_order)
    assert g_nodes.shape == o_nodes.shape
    assert np.array_equal(g_nodes, o_nodes)


def check_hits(g, o):
    assert same_hits(g, o)
    hit = o["inst"] >= 0
    rel = np.abs(g["t"][hit] - o["t"][hit]) / np.maximum(np.abs(o["t"][hit]), 1e-30)
    assert rel.max(initial=0) <= T_REL
    assert np.all(np.isinf(g["t"][~hit]))


@pytest.mark.parametrize("name", ACCEL_CASES)
def test_closest_random_rays(ctx, name):
    keep = []
    desc = scene_desc(name, keep)
    orc = ptref.OracleScene(desc)
    rays = random_rays(20000, -12, 12, seed=11)
    say(f"[gpu] closest-hit random rays {name}")
    gpu_scene(ctx, desc)
    g, gc = ctx.trace_closest(rays, counted=True)
    o, oc = orc.trace(rays, counted=True)
    check_hits(g, o)
    assert np.array_equal(gc, oc)
    b = orc.trace(rays, brute=True)
    check_hits(g, b)


@pytest.mark.parametrize("name", ["random_spheres", "cornell"])
def test_closest_primary_rays(ctx, name):
    sc = ptamd.Scene(ptamd.PT_SCENE_RANDOM_SPHERES if name == "random_spheres" else ptamd.PT_SCENE_CORNELL)
    rays = camera_rays(sc.camera, sc.info.width, sc.info.height)
    orc = ptref.OracleScene(sc.desc)
    say(f"[gpu] primary rays {name} ({len(rays)})")
    gpu_scene(ctx, sc.desc)
    g = ctx.trace_closest(rays)
    o = orc.trace(rays)
    check_hits(g, o)


def test_occluded_rule(ctx):
    sc = ptamd.Scene(ptamd.PT_SCENE_CORNELL)
    orc = ptref.OracleScene(sc.desc)
    rng = np.random.default_rng(5)
    n = 20000
    r = np.zeros(n, ptamd.RAY_DTYPE)
    r["org"] = rng.uniform([1, 1, 1], [554, 553, 554], (n, 3)).astype(np.float32)
    tgt = np.stack([rng.uniform(213, 343, n), np.full(n, 554.0), rng.uniform(227, 332, n)], 1)
    r["dir"] = (tgt - r["org"]).astype(np.float32)
    r["tmin"] = 1e-3
    r["tmax"] = np.inf
    say("[gpu] occlusion queries cornell")
    gpu_scene(ctx, sc.desc)
    g = ctx.trace_occluded(r)
    o = orc.occluded(r)
    assert np.array_equal(g, o)
    assert 0 < g.mean() < 1


def compare_paths(rgb, sig, orgb, osig, min_same=1.0):
    same = sig == osig
    assert same.mean() >= min_same, f"only {same.mean():.4f} of paths match the oracle's decisions"
    a, b = rgb[same].astype(np.float64), orgb[same].astype(np.float64)
    fin = np.isfinite(b).all(1)
    rel = np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), 1e-3)
    assert rel.max(initial=0) <= RAD_REL, f"max relative radiance error {rel.max():.3g}"
    return same.mean()


# 37x23 and 101x77: ragged path counts (1702 + 851 and 15554 + 7777 slots,
# none a multiple of 64): partial last chunks in the waves' segments and the
# identity tail of the slot permutation
@pytest.mark.parametrize("name,W,H,frames", [("cornell", 96, 96, 3), ("random_spheres", 120, 80, 3),
                                             ("cornell", 37, 23, 3), ("random_spheres", 101, 77, 3)])
@pytest.mark.parametrize("sort", [False, True])
def test_paths_match_oracle(ctx, name, W, H, frames, sort):
    cfg = ptamd.PT_SCENE_CORNELL if name == "cornell" else ptamd.PT_SCENE_RANDOM_SPHERES
    sc = ptamd.Scene(cfg, W, H)
    flags = ptamd.PT_FLAG_KEEP_PATHS | (ptamd.PT_FLAG_MATERIAL_SORT if sort else 0)
    params = sc.render_params(flags=flags, frames_per_batch=2)
    say(f"[gpu] render paths {name} {W}x{H} frames={frames} sort={sort}")
    gpu_scene(ctx, sc.desc)
    ctx.set_camera(sc.camera)
    st = ctx.render(params, 0, frames)
    rgb, sig, bounces = ctx.read_paths(W, H)
    acc, nfr = ctx.read_accum(W, H)
    orc = ptref.OracleScene(sc.desc)
    oacc, (orgb, osig, ob), ost = orc.render(sc.camera, params, 0, frames, keep_last=True)
    compare_paths(rgb, sig, orgb, osig)
    assert nfr == frames
    assert st.paths == W * H * frames
    # every path takes the oracle's decisions: the same rays are traced and
    # the accumulations and displayed bytes are identical
    assert (st.extension_rays, st.shadow_rays, st.nan_paths) == (ost[1], ost[2], ost[3])
    assert np.array_equal(acc, oacc)
    disp = ctx.read_display(W, H)
    odisp = ptref.display(oacc, frames)
    assert np.array_equal(disp[..., :3], odisp[..., :3])


def test_paths_match_oracle_light_geometry_moved(ctx):
    """The light table and the emissive geometry are independent (as in the
    reference, where the light PDFs are set up apart from the hitables):
    with the emissive rect moved half a unit below the rect the light
    samples, shadow rays meet the emitter short of the sample point or pass
    it, and the paths still match the oracle's closest-hit shadow rule."""
    W, H, frames = 64, 64, 2
    sc = ptamd.Scene(ptamd.PT_SCENE_CORNELL, W, H)
    d = sc.desc
    moved = 0
    for i in range(d.num_rects):
        if d.materials[d.rects[i].material].type == ptamd.PT_MAT_DIFFUSE_LIGHT:
            d.rects[i].k -= 0.5
            moved += 1
    assert moved == 1
    params = sc.render_params(flags=ptamd.PT_FLAG_KEEP_PATHS, frames_per_batch=2)
    say("[gpu] render paths cornell with the light geometry moved")
    gpu_scene(ctx, d)
    ctx.set_camera(sc.camera)
    ctx.render(params, 0, frames)
    rgb, sig, _ = ctx.read_paths(W, H)
    orc = ptref.OracleScene(d)
    _, (orgb, osig, _), _ = orc.render(sc.camera, params, 0, frames, keep_last=True)
    compare_paths(rgb, sig, orgb, osig)


def test_render_continues_accumulation(ctx):
    W, H = 64, 64
    sc = ptamd.Scene(ptamd.PT_SCENE_CORNELL, W, H)
    params = sc.render_params(flags=ptamd.PT_FLAG_MATERIAL_SORT)
    say("[gpu] progressive accumulation cornell")
    gpu_scene(ctx, sc.desc)
    ctx.set_camera(sc.camera)
    ctx.render(params, 0, 4)
    a4, n4 = ctx.read_accum(W, H)
    ctx.render(params, 0, 2)
    ctx.render(params, 2, 2)
    a22, n22 = ctx.read_accum(W, H)
    assert n4 == n22 == 4
    assert np.array_equal(a4, a22)


# ------------------------------------------- triangle meshes under an HDRI
# PT_SCENE_MESHES variant 0 is the full workload (8 x 2^17 triangles, 2048 x
# 1024 sky); variants 1..6 subdivide the octaspheres that many times and use
# a 256 x 128 sky.
@pytest.mark.parametrize("variant,size", [(3, (256, 128)), (0, (2048, 1024))])
def test_env_tables_bit_exact(ctx, variant, size):
    sc = ptamd.Scene(ptamd.PT_SCENE_MESHES, 64, 36, variant)
    say(f"[gpu] environment tables variant={variant}")
    ctx.set_scene(sc.desc)
    cond, marg, total = ctx.read_env_tables()
    ocond, omarg, ototal = ptref.OracleScene(sc.desc).env_tables()
    assert cond.shape == (size[1], size[0])
    assert total == ototal > 0
    assert np.array_equal(marg, omarg)
    assert np.array_equal(cond, ocond)


def test_mesh_scene_accel_and_primary_hits(ctx):
    W, H = 480, 270
    sc = ptamd.Scene(ptamd.PT_SCENE_MESHES, W, H, 0)
    orc = ptref.OracleScene(sc.desc)
    say("[gpu] mesh scene (2^20 triangles): accel build + primary rays")
    gpu_scene(ctx, sc.desc)
    g_nodes, g_order, g_morton, g_info = ctx.read_accel()
    o_nodes, o_order, o_morton, o_info = orc.accel()
    assert np.array_equal(g_info, o_info)
    assert np.array_equal(g_morton, o_morton) and np.array_equal(g_order, o_order)
    assert np.array_equal(g_nodes, o_nodes)
    rays = camera_rays(sc.camera, W, H)
    g, gc = ctx.trace_closest(rays, counted=True)
    o, oc = orc.trace(rays, counted=True, threads=16)
    check_hits(g, o)
    assert np.array_equal(gc, oc)
    assert (o["inst"] >= 0).mean() > 0.2


@pytest.mark.parametrize("variant,W,H,frames", [(4, 96, 54, 2), (0, 64, 36, 1)])
@pytest.mark.parametrize("env_mis", [True, False])
def test_mesh_paths_match_oracle(ctx, variant, W, H, frames, env_mis):
    sc = ptamd.Scene(ptamd.PT_SCENE_MESHES, W, H, variant)
    flags = ptamd.PT_FLAG_KEEP_PATHS | ptamd.PT_FLAG_MATERIAL_SORT | (ptamd.PT_FLAG_ENV_MIS if env_mis else 0)
    params = sc.render_params(flags=flags, frames_per_batch=frames)
    say(f"[gpu] render paths meshes variant={variant} {W}x{H} frames={frames} env_mis={env_mis}")
    gpu_scene(ctx, sc.desc)
    ctx.set_camera(sc.camera)
    st = ctx.render(params, 0, frames)
    rgb, sig, bounces = ctx.read_paths(W, H)
    orc = ptref.OracleScene(sc.desc)
    oacc, (orgb, osig, ob), ost = orc.render(sc.camera, params, 0, frames, keep_last=True, threads=16)
    compare_paths(rgb, sig, orgb, osig)
    assert st.paths == W * H * frames
    # the only light is the environment, sampled only with PT_FLAG_ENV_MIS
    assert (st.shadow_rays > 0) == env_mis


def uv_sphere_obj(nlat=10, nlon=20):
    """A UV sphere as OBJ text: quads with vt and vn, the upper half in
    material "textured", the lower half in "plain"."""
    lines = ["mtllib model.mtl"]
    for i in range(nlat + 1):
        th = np.pi * i / nlat
        for j in range(nlon + 1):
            ph = 2 * np.pi * j / nlon
            p = (np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph))
            lines.append("v %.6f %.6f %.6f" % p)
            lines.append("vn %.6f %.6f %.6f" % p)
            lines.append("vt %.6f %.6f" % (j / nlon, 1 - i / nlat))
    cur = None
    for i in range(nlat):
        mat = "textured" if i < nlat // 2 else "plain"
        if mat != cur:
            lines.append("usemtl " + mat)
            cur = mat
        for j in range(nlon):
            a = i * (nlon + 1) + j + 1
            b, c, d = a + nlon + 1, a + nlon + 2, a + 1
            lines.append(f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c} {d}/{d}/{d}")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("env", [True, False])
def test_obj_model_paths_match_oracle(ctx, tmp_path, env):
    """An OBJ/MTL model with a PNG diffuse map (pt_scene_build_obj), under a
    run-length-encoded Radiance environment (importance-sampled, MIS) or the
    gradient sky: per-path decisions and radiance match the oracle."""
    from asset_files import png_bytes, rgbe_bytes

    rng = np.random.default_rng(3)
    (tmp_path / "model.obj").write_text(uv_sphere_obj())
    (tmp_path / "model.mtl").write_text("newmtl textured\nKa 0 0 0\nmap_Kd checker.png\n"
                                        "newmtl plain\nKa 0.6 0.5 0.3\n")
    chk = (np.indices((16, 32)).sum(0) // 1 % 2)[..., None] * np.array([200, 60, 30]) + 40
    (tmp_path / "checker.png").write_bytes(png_bytes(chk, 2, 8))
    hdr = None
    if env:
        H, W = 16, 32
        rgbe = np.stack([rng.integers(60, 256, (H, W)) for _ in range(3)] + [rng.integers(126, 131, (H, W))], 2)
        rgbe[3:5, 10:14, 3] = 138  # a bright patch
        (tmp_path / "sky.hdr").write_bytes(rgbe_bytes(rgbe.astype(np.uint8)))
        hdr = str(tmp_path / "sky.hdr")
    sc = ptamd.Scene.from_obj("model.obj", str(tmp_path) + "/", hdr, 64, 48)
    Wd, Hd, frames = 64, 48, 2
    params = sc.render_params(flags=sc.info.flags | ptamd.PT_FLAG_KEEP_PATHS, frames_per_batch=2)
    say(f"[gpu] render paths obj model env={env}")
    gpu_scene(ctx, sc.desc)
    ctx.set_camera(sc.camera)
    ctx.render(params, 0, frames)
    rgb, sig, _ = ctx.read_paths(Wd, Hd)
    orc = ptref.OracleScene(sc.desc)
    _, (orgb, osig, _), _ = orc.render(sc.camera, params, 0, frames, keep_last=True)
    compare_paths(rgb, sig, orgb, osig)
    assert rgb.max() > 0


def look_camera(origin, W, H, vfov=40.0):
    """pinhole camera at `origin` looking at the world origin, +Y up"""
    import math

    o = np.array(origin, np.float64)
    w = o / np.linalg.norm(o)
    u = np.cross((0, 1, 0), w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    hh = math.tan(math.radians(vfov) / 2)
    hw = W / H * hh
    cam = ptamd.Camera()
    for name, val in (("lower_left", o - hw * u - hh * v - w), ("horizontal", 2 * hw * u),
                      ("vertical", 2 * hh * v), ("origin", o), ("u", u), ("v", v), ("w", w)):
        getattr(cam, name)[:] = [float(x) for x in np.float32(val)]
    cam.lens_radius, cam.time0, cam.time1 = 0.0, 0.0, 1.0
    return cam


def test_image_background_paths_match_oracle(ctx):
    """The LDR latitude-longitude background (miss.cu:51-59) and image
    textures on spheres: per-path results match the oracle."""
    s = PyScene()
    rng = np.random.default_rng(4)
    bg = np.concatenate([rng.uniform(0, 1, (8, 16, 3)), np.ones((8, 16, 1))], 2).astype(np.float32)
    tex = np.concatenate([rng.uniform(0, 1, (4, 4, 3)), np.ones((4, 4, 1))], 2).astype(np.float32)
    ti = s.texture(image=tex)
    m = s.material(ptamd.PT_MAT_LAMBERTIAN, ti)
    s.spheres += [((0, 0, 0), 1.0, m), ((2.2, 0, 0), 0.8, s.lambert((0.5, 0.5, 0.5)))]
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 2)
    s.miss = (ptamd.PT_MISS_IMAGE, (0, 0, 0), (0, 0, 0), s.texture(image=bg), 1)
    desc = s.desc()
    W, H, frames = 48, 32, 2
    cam = look_camera((1.0, 0.5, 6.0), W, H)
    params = ptamd.RenderParams(W, H, 8, 1, 3, 2, ptamd.PT_FLAG_KEEP_PATHS, 0)
    say("[gpu] render paths image background")
    gpu_scene(ctx, desc)
    ctx.set_camera(cam)
    ctx.render(params, 0, frames)
    rgb, sig, _ = ctx.read_paths(W, H)
    orc = ptref.OracleScene(desc)
    _, (orgb, osig, _), _ = orc.render(cam, params, 0, frames, keep_last=True)
    compare_paths(rgb, sig, orgb, osig)


def test_procedural_textures_paths_match_oracle(ctx):
    """Checker (checkered_texture.cu, one child an image texture sampled at
    the sphere's uv), Perlin noise (noise_texture.cu) and gradient
    (gradient_texture.cu) textures under Lambertian, metal and
    Ashikhmin-Shirley materials: per-path results match the oracle; a
    checker of a checker is refused."""
    rng = np.random.default_rng(8)
    s = PyScene()
    img = np.concatenate([rng.uniform(0, 1, (6, 12, 3)), np.ones((6, 12, 1))], 2).astype(np.float32)
    im = s.texture(image=img)
    red, white = s.texture((0.8, 0.1, 0.1)), s.texture((0.9, 0.9, 0.9))
    chk = s.checker(im, white)
    vec = rng.normal(size=(256, 3))
    vec /= np.linalg.norm(vec, axis=1, keepdims=True)
    perms = np.stack([rng.permutation(256) for _ in range(3)])
    noise = s.noise(4.0, vec.astype(np.float32), perms)
    grad = s.gradient((0.9, 0.2, 0.1), (0.1, 0.8, 0.2), (0.2, 0.3, 0.9))
    chk2 = s.checker(red, grad)
    s.spheres += [((0, -100.5, 0), 100.0, s.material(ptamd.PT_MAT_LAMBERTIAN, chk)),
                  ((0, 0, 0), 0.5, s.material(ptamd.PT_MAT_METAL, noise, params=(0.1,))),
                  ((1.1, 0, 0), 0.5, s.material(ptamd.PT_MAT_LAMBERTIAN, grad)),
                  ((-1.1, 0, 0), 0.5, s.material(ptamd.PT_MAT_ASHIKHMIN_SHIRLEY, chk2, noise, (10.0, 100.0)))]
    s.instance(ptamd.PT_GEOM_SPHERES, 0, len(s.spheres))
    s.miss = (ptamd.PT_MISS_GRADIENT, (1, 1, 1), (0.5, 0.7, 1.0), -1, 1)
    desc = s.desc()
    W, H, frames = 64, 40, 2
    cam = look_camera((0.5, 1.2, 4.0), W, H)
    params = ptamd.RenderParams(W, H, 8, 1, 3, 2, ptamd.PT_FLAG_KEEP_PATHS | ptamd.PT_FLAG_MATERIAL_SORT, 0)
    say("[gpu] render paths procedural textures")
    gpu_scene(ctx, desc)
    ctx.set_camera(cam)
    ctx.render(params, 0, frames)
    rgb, sig, _ = ctx.read_paths(W, H)
    orc = ptref.OracleScene(desc)
    _, (orgb, osig, _), _ = orc.render(cam, params, 0, frames, keep_last=True)
    compare_paths(rgb, sig, orgb, osig)
    assert rgb.max() > 0
    bad = PyScene()
    a, b = bad.texture((1, 1, 1)), bad.texture((0, 0, 0))
    nested = bad.checker(bad.checker(a, b), b)
    bad.spheres.append(((0, 0, 0), 1.0, bad.material(ptamd.PT_MAT_LAMBERTIAN, nested)))
    bad.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    with pytest.raises(ptamd.PtError):
        ctx.set_scene(bad.desc())


def test_moving_spheres_paths_match_oracle(ctx):
    """scenes.hpp:95-186 (checkered ground, noise-textured metal sphere,
    thin-lens camera, rect light with MIS under a dark sky): per-path
    results match the oracle."""
    sc = ptamd.Scene(ptamd.PT_SCENE_MOVING_SPHERES, 64, 43)
    W, H, frames = 64, 43, 2
    params = sc.render_params(flags=sc.info.flags | ptamd.PT_FLAG_KEEP_PATHS, frames_per_batch=2)
    say("[gpu] render paths moving spheres")
    gpu_scene(ctx, sc.desc)
    ctx.set_camera(sc.camera)
    ctx.render(params, 0, frames)
    rgb, sig, _ = ctx.read_paths(W, H)
    orc = ptref.OracleScene(sc.desc)
    _, (orgb, osig, _), _ = orc.render(sc.camera, params, 0, frames, keep_last=True)
    compare_paths(rgb, sig, orgb, osig)
    assert rgb.max() > 0


def media_scene():
    """Constant-density media: a sphere medium around a Lambertian ball, a
    box medium under a rotation and a scale, a thin medium around the whole
    scene; a normal-shaded mesh with vertex normals; a rect light."""
    s = PyScene()
    blue = s.material(ptamd.PT_MAT_ISOTROPIC, s.texture((0.2, 0.4, 0.9)))
    white = s.material(ptamd.PT_MAT_ISOTROPIC, s.texture((1, 1, 1)))
    s.volumes += [(ptamd.PT_VOLUME_SPHERE, (0, 1, 0), (1.2, 0, 0), 0.8, blue),
                  (ptamd.PT_VOLUME_SPHERE, (0, 0, 0), (50, 0, 0), 0.01, white),
                  (ptamd.PT_VOLUME_BOX, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 1.5, blue)]
    s.instance(ptamd.PT_GEOM_VOLUMES, 0, 2)
    c, sn = np.cos(0.5), np.sin(0.5)
    M = np.array([[c * 1.5, 0, sn, 2.2], [0, 1.0, 0, 0.6], [-sn * 1.5, 0, c, 0.3]])
    s.instance(ptamd.PT_GEOM_VOLUMES, 2, 1, M)
    s.spheres += [((0, 1, 0), 0.5, s.lambert((0.8, 0.3, 0.2))), ((0, -100, 0), 100.0, s.lambert((0.5, 0.5, 0.5)))]
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 2)
    v, f = icosphere(2, 0.6, (-2.0, 0.7, 0.5))
    nrm = (v - np.float32([-2.0, 0.7, 0.5])) * np.float32(-1)  # inward vertex normals
    s.meshes.append((v, f, nrm, None, s.material(ptamd.PT_MAT_NORMAL, -1, params=(0.0,))))
    s.instance(ptamd.PT_GEOM_MESH, 0, 0)
    light = s.material(ptamd.PT_MAT_DIFFUSE_LIGHT, s.texture((6, 6, 6)))
    s.rects.append((ptamd.PT_AXIS_Y, 1, -1.5, 1.5, -1.5, 1.5, 4.0, light))
    s.instance(ptamd.PT_GEOM_RECTS, 0, 1)
    s.lights.append({"type": ptamd.PT_LIGHT_RECT, "axis": ptamd.PT_AXIS_Y, "rect": (-1.5, 1.5, -1.5, 1.5, 4.0),
                     "emission": (6, 6, 6)})
    s.miss = (ptamd.PT_MISS_CONSTANT, (0.05, 0.05, 0.08), (0, 0, 0), -1, 1)
    return s


def test_volume_queries_match_oracle(ctx):
    """Closest-hit and occlusion queries through media (free flights drawn
    from the ray index): the same hits and distances as the oracle's BVH
    and brute force."""
    s = media_scene()
    desc = s.desc()
    orc = ptref.OracleScene(desc)
    rays = random_rays(20000, -3, 3, seed=12)
    say("[gpu] closest-hit queries through media")
    gpu_scene(ctx, desc)
    g, gc = ctx.trace_closest(rays, counted=True)
    o, oc = orc.trace(rays, counted=True)
    check_hits(g, o)
    assert np.array_equal(gc, oc)
    check_hits(g, orc.trace(rays, brute=True))
    vol_hits = (g["inst"] <= 1) & (g["inst"] >= 0)
    assert vol_hits.sum() > 1000
    assert np.array_equal(ctx.trace_occluded(rays), orc.occluded(rays))


def test_media_paths_match_oracle(ctx):
    """Isotropic scattering in sphere and box media (light sampling without
    the hemisphere test, shadow rays through media), the normal shader on a
    mesh with vertex normals: per-path results match the oracle."""
    s = media_scene()
    desc = s.desc()
    W, H, frames = 64, 48, 2
    cam = look_camera((0.5, 2.0, 6.0), W, H, vfov=50.0)
    for flags in (ptamd.PT_FLAG_KEEP_PATHS, ptamd.PT_FLAG_KEEP_PATHS | ptamd.PT_FLAG_MATERIAL_SORT):
        params = ptamd.RenderParams(W, H, 12, 1, 3, 2, flags, 0)
        say(f"[gpu] render paths media flags={flags}")
        gpu_scene(ctx, desc)
        ctx.set_camera(cam)
        st = ctx.render(params, 0, frames)
        rgb, sig, _ = ctx.read_paths(W, H)
        orc = ptref.OracleScene(desc)
        _, (orgb, osig, _), ost = orc.render(cam, params, 0, frames, keep_last=True)
        compare_paths(rgb, sig, orgb, osig)
        assert st.shadow_rays == ost[2] and st.extension_rays == ost[1]
        assert rgb.max() > 0


def test_next_week_paths_match_oracle(ctx):
    """scenes.hpp:263-375 (ground boxes, media, Perlin sphere, 1000-sphere
    group under a rotation and translation): per-path results match."""
    sc = ptamd.Scene(ptamd.PT_SCENE_NEXT_WEEK, 64, 64)
    W, H, frames = 64, 64, 2
    params = sc.render_params(flags=sc.info.flags | ptamd.PT_FLAG_KEEP_PATHS, frames_per_batch=2)
    say("[gpu] render paths next week")
    gpu_scene(ctx, sc.desc)
    ctx.set_camera(sc.camera)
    ctx.render(params, 0, frames)
    rgb, sig, _ = ctx.read_paths(W, H)
    orc = ptref.OracleScene(sc.desc)
    _, (orgb, osig, _), _ = orc.render(sc.camera, params, 0, frames, keep_last=True)
    compare_paths(rgb, sig, orgb, osig)
    assert rgb.max() > 0


def shapes_scene():
    """Open cylinders (one rotated and scaled by an instance), moving
    spheres swept over the shutter, a ground sphere and a sphere light."""
    s = PyScene()
    red, white = s.lambert((0.8, 0.2, 0.1)), s.lambert((0.7, 0.7, 0.7))
    metal = s.material(ptamd.PT_MAT_METAL, s.texture((0.8, 0.8, 0.9)), params=(0.05,))
    s.cylinders += [((0.0, -0.5, 0.0), 1.5, 0.6, red), ((0.0, 0.0, 0.0), 1.0, 0.5, metal)]
    s.instance(ptamd.PT_GEOM_CYLINDERS, 0, 1)
    c, sn = np.cos(1.2), np.sin(1.2)
    s.instance(ptamd.PT_GEOM_CYLINDERS, 1, 1, np.array([[c, -sn, 0, 1.8], [sn, c, 0, 0.2], [0, 0, 1, -0.5]]))
    s.moving += [((-2.2, 0.0, 0.8), (-1.2, 0.6, 0.8), 0.4, 0.0, 1.0, white),
                 ((1.5, 1.2, 1.0), (1.5, 0.2, 1.0), 0.3, 0.0, 1.0, metal)]
    s.instance(ptamd.PT_GEOM_MOVING_SPHERES, 0, 2)
    s.spheres += [((0, -100.5, 0), 100.0, white), ((0, 5, 2), 0.8, s.material(ptamd.PT_MAT_DIFFUSE_LIGHT,
                                                                                 s.texture((8, 8, 8))))]
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 2)
    s.lights.append({"type": ptamd.PT_LIGHT_SPHERE, "center": (0, 5, 2), "radius": 0.8, "emission": (8, 8, 8)})
    return s


def test_shapes_queries_and_paths_match_oracle(ctx):
    """Closest-hit queries (moving spheres at time 0) and per-path results
    with each path's time drawn in the shutter [0, 1]: GPU == oracle."""
    s = shapes_scene()
    desc = s.desc()
    orc = ptref.OracleScene(desc)
    rays = random_rays(20000, -3, 3, seed=13)
    say("[gpu] closest-hit queries cylinders and moving spheres")
    gpu_scene(ctx, desc)
    g, gc = ctx.trace_closest(rays, counted=True)
    o, oc = orc.trace(rays, counted=True)
    check_hits(g, o)
    assert np.array_equal(gc, oc)
    assert ((g["inst"] >= 0) & (g["inst"] <= 2)).sum() > 500
    W, H, frames = 64, 48, 2
    cam = look_camera((0.3, 1.5, 6.0), W, H, vfov=45.0)
    cam.time0, cam.time1 = 0.0, 1.0
    for flags in (ptamd.PT_FLAG_KEEP_PATHS, ptamd.PT_FLAG_KEEP_PATHS | ptamd.PT_FLAG_MATERIAL_SORT):
        params = ptamd.RenderParams(W, H, 10, 1, 3, 2, flags, 0)
        say(f"[gpu] render paths cylinders and moving spheres flags={flags}")
        ctx.set_camera(cam)
        ctx.render(params, 0, frames)
        rgb, sig, _ = ctx.read_paths(W, H)
        _, (orgb, osig, _), _ = orc.render(cam, params, 0, frames, keep_last=True)
        compare_paths(rgb, sig, orgb, osig)
        assert rgb.max() > 0


def test_many_instances_match_oracle(ctx):
    """Two-level instancing at scale (hitables.hpp:501-515, mesh.hpp:317-332):
    5000 rotated and scaled instances of one mesh structure plus a
    1000-sphere group under one transform.  The top-level structure is
    bit-exact, closest hits match the oracle's traversal, and per-path
    results match."""
    s = mesh_scene(1, 5000, 21)
    rng = np.random.default_rng(22)
    m = s.lambert((0.3, 0.6, 0.9))
    s.spheres += [(tuple(rng.uniform(0, 8, 3)), 0.3, m) for _ in range(1000)]
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1000, np.array([[0.8, -0.6, 0, -4], [0.6, 0.8, 0, 2], [0, 0, 1, 1]]))
    desc = s.desc()
    say("[gpu] accel build and queries, 5001 instances")
    gpu_scene(ctx, desc)
    g_nodes, g_order, g_morton, g_info = ctx.read_accel()
    orc = ptref.OracleScene(desc)
    o_nodes, o_order, o_morton, o_info = orc.accel()
    assert np.array_equal(g_info, o_info) and np.array_equal(g_order, o_order)
    assert np.array_equal(g_nodes, o_nodes)
    rays = random_rays(20000, -12, 12, seed=23)
    g, gc = ctx.trace_closest(rays, counted=True)
    o, oc = orc.trace(rays, counted=True)
    check_hits(g, o)
    assert np.array_equal(gc, oc)
    assert (g["inst"] == 5000).sum() > 50 and (g["inst"] >= 0).mean() > 0.5
    W, H, frames = 48, 32, 2
    cam = look_camera((0.0, 3.0, 24.0), W, H, vfov=50.0)
    params = ptamd.RenderParams(W, H, 6, 1, 3, 2, ptamd.PT_FLAG_KEEP_PATHS, 0)
    say("[gpu] render paths 5001 instances")
    ctx.set_camera(cam)
    ctx.render(params, 0, frames)
    rgb, sig, _ = ctx.read_paths(W, H)
    _, (orgb, osig, _), _ = orc.render(cam, params, 0, frames, keep_last=True)
    compare_paths(rgb, sig, orgb, osig)


def test_extended_kinds_rejected(ctx):
    """Malformed media, cylinders and moving spheres are refused by
    pt_set_scene with an error; a well-formed media scene renders, and the
    radiance cache refuses it."""
    def scene(**kw):
        s = PyScene()
        m = s.lambert((0.5, 0.5, 0.5))
        if "volume" in kw:
            s.volumes.append(kw["volume"] + (m,))
            s.instance(ptamd.PT_GEOM_VOLUMES, 0, 1)
        if "moving" in kw:
            s.moving.append(kw["moving"] + (m,))
            s.instance(ptamd.PT_GEOM_MOVING_SPHERES, 0, 1)
        if "cyl_mat" in kw:
            s.cylinders.append(((0, 0, 0), 1.0, 1.0, kw["cyl_mat"]))
            s.instance(ptamd.PT_GEOM_CYLINDERS, 0, 1)
        if "mat" in kw:
            s.materials[m] = kw["mat"]
            s.spheres.append(((0, 0, 0), 1.0, m))
            s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
        return s
    bad = [scene(volume=(ptamd.PT_VOLUME_SPHERE, (0, 0, 0), (1, 0, 0), 0.0)),
           scene(volume=(ptamd.PT_VOLUME_SPHERE, (0, 0, 0), (1, 0, 0), float("inf"))),
           scene(volume=(7, (0, 0, 0), (1, 0, 0), 1.0)),
           scene(moving=((0, 0, 0), (1, 0, 0), 0.5, 1.0, 1.0)),
           scene(cyl_mat=5),
           scene(mat=(ptamd.PT_MAT_LAMBERTIAN, -1, -1, (0, 0, 0, 0)))]
    say("[gpu] malformed extended-kind scenes")
    for s in bad:
        with pytest.raises(ptamd.PtError):
            ctx.set_scene(s.desc())
    good = scene(volume=(ptamd.PT_VOLUME_BOX, (0, 0, 0), (1, 1, 1), 0.5))
    sc = ptamd.Scene(ptamd.PT_SCENE_CORNELL, 16, 16)
    gpu_scene(ctx, good.desc())
    ctx.set_camera(sc.camera)
    ctx.render(ptamd.RenderParams(16, 16, 4, 1, 3, 1, 0, 0), 0, 1)
    own = ptamd.Context(0)
    try:
        gpu_scene(own, good.desc())
        own.set_camera(sc.camera)
        own.cache_init(9)
        with pytest.raises(ptamd.PtError):
            own.render(ptamd.RenderParams(16, 16, 4, 1, 3, 1, ptamd.PT_FLAG_RADIANCE_CACHE, 0), 0, 1)
    finally:
        own.close()


def test_vector_texture_paths_match_oracle(ctx):
    """vector.cu: a mesh whose faces pick a constant, an image or a checker
    (or black, out of range) through per-face indices, and a sphere (index
    0) under the same Lambertian: per-path results match the oracle."""
    rng = np.random.default_rng(31)
    s = PyScene()
    red = s.texture((0.8, 0.1, 0.1))
    img = np.concatenate([rng.uniform(0, 1, (8, 8, 3)), np.ones((8, 8, 1))], 2).astype(np.float32)
    im = s.texture(image=img)
    chk = s.checker(s.texture((0.1, 0.1, 0.8)), s.texture((0.9, 0.9, 0.9)))
    s.textures.append((ptamd.PT_TEX_VECTOR, (0, 0, 0), None, (3, 0), np.float32([red, im, chk])))
    vec = len(s.textures) - 1
    lam = s.material(ptamd.PT_MAT_LAMBERTIAN, vec)
    v, f = icosphere(2, 1.0, (0.0, 0.0, 0.0))
    uv = np.stack([v[:, 0] * 0.5 + 0.5, v[:, 1] * 0.5 + 0.5], 1).astype(np.float32)
    s.meshes.append((v, f, None, uv, lam, np.arange(len(f), dtype=np.int32) % 4))
    s.instance(ptamd.PT_GEOM_MESH, 0, 0)
    s.spheres += [((2.2, 0, 0), 0.7, lam), ((0, -100.8, 0), 100.0, s.lambert((0.5, 0.5, 0.5)))]
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 2)
    desc = s.desc()
    W, H, frames = 64, 40, 2
    cam = look_camera((0.8, 1.0, 5.0), W, H)
    params = ptamd.RenderParams(W, H, 6, 1, 3, 2, ptamd.PT_FLAG_KEEP_PATHS, 0)
    say("[gpu] render paths vector texture")
    gpu_scene(ctx, desc)
    ctx.set_camera(cam)
    ctx.render(params, 0, frames)
    rgb, sig, _ = ctx.read_paths(W, H)
    orc = ptref.OracleScene(desc)
    _, (orgb, osig, _), _ = orc.render(cam, params, 0, frames, keep_last=True)
    compare_paths(rgb, sig, orgb, osig)
    assert rgb.max() > 0
    bad = PyScene()
    a = bad.texture((1, 1, 1))
    bad.textures.append((ptamd.PT_TEX_VECTOR, (0, 0, 0), None, (2, 0), np.float32([a, 7])))
    bad.spheres.append(((0, 0, 0), 1.0, bad.material(ptamd.PT_MAT_LAMBERTIAN, len(bad.textures) - 1)))
    bad.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    with pytest.raises(ptamd.PtError):
        ctx.set_scene(bad.desc())
