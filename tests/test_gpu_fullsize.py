"""GPU parity at the benchmark configurations, through the C ABI.

  * Cornell box + MIS, 512x512, 64 spp, depth 16, RR from 3, all 64 frames in
    one batch of 16 Mi paths (the benchmarked launch shape), and the random
    spheres at 400x267, 16 spp, depth 50: the accumulation equals the
    oracle's bit for bit (every path takes the oracle's decisions), so do the
    displayed bytes and the traced ray counts;
  * the 2^20-triangle mesh scene under the importance-sampled HDRI at
    1920x1080: the full frame is rendered, and the stratified 1/64 pixel
    subset equals the oracle's per-pixel sums bit for bit;
  * relMSE against the converged FP64 references of tests/golden/fp64 is
    reported; being bit-identical, it is the oracle's own error;
  * the NaN counter of the render statistics equals the oracle's count on a
    scene whose Ashikhmin-Shirley frame is NaN (shading normal along +Y).
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import ptamd
import ptref
from helpers import PyScene
from ptamd.metrics import accum_at, relmse, stratified_pixels

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "fp64"
CPU_THREADS = 16


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ptamd.Context(0)
    yield c
    c.close()


def ref_relmse(name, img_at_pix):
    z = np.load(GOLDEN / f"{name}.npz")
    return relmse(img_at_pix, z["mean"])


@pytest.mark.parametrize("name,cfg", [("c1_cornell", ptamd.PT_SCENE_CORNELL),
                                      ("c0_random_spheres", ptamd.PT_SCENE_RANDOM_SPHERES)])
def test_full_size_render_is_bit_exact(ctx, name, cfg):
    sc = ptamd.Scene(cfg)
    W, H, spp = sc.info.width, sc.info.height, sc.info.spp
    say(f"[gpu] full-size render {name} {W}x{H} {spp} spp")
    prm = sc.render_params(frames_per_batch=spp)
    ctx.set_scene(sc.desc)
    ctx.build_accel()
    ctx.set_camera(sc.camera)
    st = ctx.render(prm, 0, spp)
    acc, frames = ctx.read_accum(W, H)
    disp = ctx.read_display(W, H)
    assert frames == spp and st.paths == W * H * spp
    say(f"[cpu] oracle {name}")
    orc = ptref.OracleScene(sc.desc)
    oacc, _, ost = orc.render(sc.camera, prm, 0, spp, threads=CPU_THREADS)
    assert (st.extension_rays, st.shadow_rays, st.nan_paths) == (ost[1], ost[2], ost[3])
    assert np.array_equal(acc, oacc)
    assert np.array_equal(disp[..., :3], ptref.display(oacc, spp)[..., :3])
    pix = stratified_pixels(W, H)
    e = ref_relmse(name, accum_at(acc, pix, W, H) / spp)
    say(f"[gpu] {name}: relMSE vs FP64 reference {e:.5f}")
    assert np.isfinite(e) and e < 0.5


def test_mesh_scene_full_frame_subset_is_bit_exact(ctx):
    sc = ptamd.Scene(ptamd.PT_SCENE_MESHES)
    W, H = sc.info.width, sc.info.height
    frames = 4
    say(f"[gpu] mesh scene {W}x{H}, {frames} frames")
    prm = sc.render_params(frames_per_batch=frames)
    ctx.set_scene(sc.desc)
    ctx.build_accel()
    ctx.set_camera(sc.camera)
    st = ctx.render(prm, 0, frames)
    acc, _ = ctx.read_accum(W, H)
    assert st.paths == W * H * frames
    pix = stratified_pixels(W, H)
    say(f"[cpu] oracle on {len(pix)} pixels")
    orc = ptref.OracleScene(sc.desc)
    sums, _ = orc.render_pixels(sc.camera, prm, 0, frames, pix, threads=CPU_THREADS)
    got = accum_at(acc, pix, W, H)
    assert np.array_equal(got, sums)
    e = ref_relmse("c2_meshes", got / frames) if (GOLDEN / "c2_meshes.npz").exists() else float("nan")
    say(f"[gpu] c2_meshes at {frames} spp: relMSE vs FP64 reference {e:.4f}")


def nan_scene():
    """An Ashikhmin-Shirley floor facing +Y under an importance-sampled
    environment: the floor's tangent frame normalize(cross(N, Y)) is NaN
    (decision D9), so its continuation ray is NaN, the environment it misses
    into returns NaN unclamped (environment MIS), and the path's color is
    NaN."""
    s = PyScene()
    img = np.full((16, 32, 4), 0.5, np.float32)
    img[4, 9, :3] = 20.0
    s.miss = (ptamd.PT_MISS_HDR, (0, 0, 0), (0, 0, 0), s.texture(image=img), 1)
    s.lights.append({"type": ptamd.PT_LIGHT_ENV, "emission": (1, 1, 1)})
    t = s.texture((0.6, 0.5, 0.4))
    ts = s.texture((0.3, 0.3, 0.3))
    asm = s.material(ptamd.PT_MAT_ASHIKHMIN_SHIRLEY, t, ts, (100.0, 100.0))
    s.rects.append((ptamd.PT_AXIS_Y, 0, -5.0, 5.0, -5.0, 5.0, 0.0, asm))
    s.spheres.append(((0, 1, 0), 1.0, s.lambert((0.7, 0.7, 0.7))))
    s.instance(ptamd.PT_GEOM_RECTS, 0, 1)
    s.instance(ptamd.PT_GEOM_SPHERES, 0, 1)
    return s


def test_nan_paths_are_counted_as_the_oracle_counts(ctx):
    s = nan_scene()
    desc = s.desc()
    cam = ptamd.Camera()
    cam.origin[:] = [0, 3, 8]
    cam.lower_left[:] = [-2, 0, 4]
    cam.horizontal[:] = [4, 0, 0]
    cam.vertical[:] = [0, 3, 0]
    cam.time1 = 1.0
    W, H, frames = 64, 48, 4
    prm = ptamd.RenderParams(W, H, 8, 1, 3, frames, ptamd.PT_FLAG_ENV_MIS, 0)
    ctx.set_scene(desc)
    ctx.build_accel()
    ctx.set_camera(cam)
    st = ctx.render(prm, 0, frames)
    acc, _ = ctx.read_accum(W, H)
    oacc, _, ost = ptref.OracleScene(desc).render(cam, prm, 0, frames)
    assert st.nan_paths == ost[3] > W * H * frames // 10
    assert np.array_equal(acc, oacc)
    assert np.all(np.isfinite(acc))
