"""A batch of scenes in one wavefront (pt_render_batch): every context's
accumulation equals, bit for bit, its own pt_render of the same frames, and
the batch's ray counts are the sums of the single renders'."""
import sys

import numpy as np
import pytest

import ptamd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make(cfg, W, H, variant=0):
    sc = ptamd.Scene(cfg, W, H, variant)
    ctx = ptamd.Context(0)
    ctx.set_scene(sc.desc)
    ctx.build_accel()
    ctx.set_camera(sc.camera)
    return sc, ctx


@pytest.fixture(scope="module")
def scenes():
    # Cornell (depth 16, material sort), random spheres (depth 50, gradient
    # sky), coarse meshes under the HDRI with environment MIS (depth 12)
    out = [make(ptamd.PT_SCENE_CORNELL, 64, 64), make(ptamd.PT_SCENE_RANDOM_SPHERES, 80, 54),
           make(ptamd.PT_SCENE_MESHES, 64, 36, 6), make(ptamd.PT_SCENE_CORNELL, 40, 24)]
    yield out
    for _, ctx in out:
        ctx.close()


def single(sc, ctx, frames):
    st = ctx.render(sc.render_params(), 0, frames)
    acc, n = ctx.read_accum(sc.info.width, sc.info.height)
    return acc, n, st


def test_batch_equals_single_renders(scenes):
    frames = [5, 3, 4, 7]
    alone = []
    for (sc, ctx), f in zip(scenes, frames):
        say(f"[gpu] single render {sc.info.width}x{sc.info.height} x {f}")
        alone.append(single(sc, ctx, f))
    say("[gpu] batch of 4 scenes")
    st = ptamd.render_batch([c for _, c in scenes], [s.render_params() for s, _ in scenes], frames)
    for (sc, ctx), (acc, n, _), f in zip(scenes, alone, frames):
        got, m = ctx.read_accum(sc.info.width, sc.info.height)
        assert m == n == f
        assert np.array_equal(got, acc)
    assert st.paths == sum(s.info.width * s.info.height * f for (s, _), f in zip(scenes, frames))
    assert st.extension_rays == sum(a[2].extension_rays for a in alone)
    assert st.shadow_rays == sum(a[2].shadow_rays for a in alone)
    assert st.exhausted_paths == sum(a[2].exhausted_paths for a in alone)
    # the contexts render on their own afterwards as before
    sc, ctx = scenes[1]
    a2, _, _ = single(sc, ctx, frames[1])
    assert np.array_equal(a2, alone[1][0])


def test_batch_with_media_and_moving_spheres_equals_single_renders():
    """Scenes with constant media, isotropic phase functions, moving spheres
    and noise textures (The Next Week, the moving-spheres scene) batched with
    the Cornell box: each accumulation equals its scene's own render bit for
    bit (the media's free flights hash only the RNG state's low 24 bits, so
    the scene index in the top byte leaves them unchanged)."""
    made = [make(ptamd.PT_SCENE_NEXT_WEEK, 48, 48), make(ptamd.PT_SCENE_MOVING_SPHERES, 40, 27),
            make(ptamd.PT_SCENE_CORNELL, 32, 32)]
    try:
        frames = [3, 2, 2]
        alone = []
        for (sc, ctx), f in zip(made, frames):
            say(f"[gpu] single render {sc.info.width}x{sc.info.height} x {f}")
            alone.append(single(sc, ctx, f))
        say("[gpu] batch: next week, moving spheres, cornell")
        st = ptamd.render_batch([c for _, c in made], [s.render_params() for s, _ in made], frames)
        for (sc, ctx), (acc, n, _), f in zip(made, alone, frames):
            got, m = ctx.read_accum(sc.info.width, sc.info.height)
            assert m == n == f
            assert np.array_equal(got, acc)
        assert st.extension_rays == sum(a[2].extension_rays for a in alone)
        assert st.shadow_rays == sum(a[2].shadow_rays for a in alone)
    finally:
        for _, ctx in made:
            ctx.close()


def test_batch_split_across_path_pools(scenes):
    """512 x 512 x 70 frames exceed one 16 Mi-path batch: the work list is cut
    inside the large scene's frames and the small scenes follow."""
    big_sc, big = make(ptamd.PT_SCENE_CORNELL, 512, 512)
    try:
        ctxs = [big, scenes[0][1], scenes[2][1]]
        scs = [big_sc, scenes[0][0], scenes[2][0]]
        frames = [70, 2, 3]
        alone = [single(s, c, f)[0] for s, c, f in zip(scs, ctxs, frames)]
        say("[gpu] batch 512x512x70 + 2 small scenes")
        ptamd.render_batch(ctxs, [s.render_params() for s in scs], frames, stats=False)
        for s, c, a in zip(scs, ctxs, alone):
            got, _ = c.read_accum(s.info.width, s.info.height)
            assert np.array_equal(got, a)
    finally:
        big.close()


def test_batch_rejects_bad_input(scenes):
    sc, ctx = scenes[0]
    p = sc.render_params()
    with pytest.raises(ptamd.PtError):
        ptamd.render_batch([ctx, ctx], [p, p], [1, 1])
    with pytest.raises(ptamd.PtError):
        ptamd.render_batch([ctx], [sc.render_params(flags=ptamd.PT_FLAG_RADIANCE_CACHE)], [1])
