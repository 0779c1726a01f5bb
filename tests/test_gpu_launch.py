"""Launch-shape machinery on the GPU: batches replayed from a captured
hipGraph and the autotuned launch shape give images bit-identical to the
directly launched render (every shape traces the same paths)."""
import sys

import numpy as np
import pytest

import ptamd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ptamd.Context(0)
    yield c
    c.close()


def setup(ctx, cfg, W, H, variant=0):
    sc = ptamd.Scene(cfg, W, H, variant)
    ctx.set_scene(sc.desc)
    ctx.build_accel()
    ctx.set_camera(sc.camera)
    return sc


def image(ctx, params, f0, n):
    st = ctx.render(params, f0, n)
    acc, frames = ctx.read_accum(params.width, params.height)
    return acc, st


# 45x31: a ragged path count (2790 slots per two-frame batch)
@pytest.mark.parametrize("cfg,W,H", [(ptamd.PT_SCENE_CORNELL, 96, 64), (ptamd.PT_SCENE_MESHES, 96, 64),
                                     (ptamd.PT_SCENE_CORNELL, 45, 31)])
def test_graph_replay_is_bit_identical(ctx, cfg, W, H):
    sc = setup(ctx, cfg, W, H, 3 if cfg == ptamd.PT_SCENE_MESHES else 0)
    say(f"[gpu] graph replay scene {cfg} {W}x{H}")
    base = sc.render_params(frames_per_batch=2)
    ref, _ = image(ctx, base, 0, 7)
    g = sc.render_params(flags=sc.info.flags | ptamd.PT_FLAG_GRAPH, frames_per_batch=2)
    a, st = image(ctx, g, 0, 7)      # 3 graph replays and a direct 1-frame tail
    assert np.array_equal(a, ref)
    assert st.extend_ms == 0  # no per-kernel events inside graphs
    b, _ = image(ctx, g, 0, 7)       # replays of the cached graph
    assert np.array_equal(b, ref)
    # progressive continuation through graphs
    image(ctx, g, 0, 4)
    c2, _ = image(ctx, g, 4, 3)
    assert np.array_equal(c2, ref)


def test_autotune_picks_a_shape_with_identical_results(ctx):
    W = H = 128
    sc = setup(ctx, ptamd.PT_SCENE_CORNELL, W, H)
    say("[gpu] autotune cornell 128x128")
    p = sc.render_params()
    ref, _ = image(ctx, p, 0, 128)
    r = ctx.autotune(p, 128)
    say(f"[gpu] tuned: frames_per_batch={r.frames_per_batch} sort={r.material_sort} graph={r.graph} "
        f"{r.ms_per_frame:.3f} ms/frame over {r.candidates} candidates")
    assert r.candidates >= 4 and r.frames_per_batch >= 1 and r.ms_per_frame > 0
    t = sc.render_params(flags=p.flags | ptamd.PT_FLAG_TUNED)
    a, _ = image(ctx, t, 0, 128)
    assert np.array_equal(a, ref)


DRIVER = ptamd.REPO / "tests" / "_bin" / "facade_driver"


@pytest.mark.parametrize("scene,cfg,W,H,frames", [("Cornell", ptamd.PT_SCENE_CORNELL, 96, 80, 4),
                                                  ("RandomSpheres", ptamd.PT_SCENE_RANDOM_SPHERES, 90, 60, 3)])
def test_facade_launch_loop_equals_render(tmp_path, scene, cfg, W, H, frames):
    """The reference's main.cpp pattern through the host scene API: the
    scene function, acc_buffer / display_buffer bound as context variables,
    an empty launch that builds, then context["frame"] and
    Context::launch(0, W, H) per frame, accBuffer->map() and
    displayBuffer->map() (tests/facade_driver.cpp).  The mapped buffers
    equal one pt_render of the same frames."""
    import subprocess

    if not DRIVER.exists():
        pytest.fail("tests/_bin/facade_driver is not built (run __graft_entry__.build())")
    out = tmp_path / "maps.bin"
    say(f"[gpu] facade driver {scene} {W}x{H} x {frames}")
    subprocess.run([str(DRIVER), scene, str(W), str(H), str(frames), str(out)], check=True, timeout=120)
    raw = out.read_bytes()
    acc_map = np.frombuffer(raw[:W * H * 16], np.float32).reshape(H, W, 4)
    disp_map = np.frombuffer(raw[W * H * 16:], np.uint8).reshape(H, W, 4)
    sc = ptamd.Scene(cfg, W, H)
    with ptamd.Context(0) as ctx:
        ctx.set_scene(sc.desc)
        ctx.build_accel()
        ctx.set_camera(sc.camera)
        ctx.render(sc.render_params(), 0, frames)
        acc, n = ctx.read_accum(W, H)
        disp = ctx.read_display(W, H)
    assert n == frames
    assert np.array_equal(acc_map[..., :3], acc.astype(np.float32))
    assert np.all(acc_map[..., 3] == frames)
    assert np.array_equal(disp_map, disp)
