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
    the Cornell box and the mesh scene under the importance-sampled HDRI
    (environment MIS): each accumulation equals its scene's own render bit
    for bit (the media's free flights hash only the RNG state's low 24 bits,
    so the scene index in the top byte leaves them unchanged).  The batch
    runs the extended-kind shading with environment sampling; run a second
    time with traversal counting, its ray, exhaustion and fetch counts are
    the sums of the single renders'."""
    made = [make(ptamd.PT_SCENE_NEXT_WEEK, 48, 48), make(ptamd.PT_SCENE_MOVING_SPHERES, 40, 27),
            make(ptamd.PT_SCENE_CORNELL, 32, 32), make(ptamd.PT_SCENE_MESHES, 64, 36, 6)]
    try:
        frames = [3, 2, 2, 3]
        assert made[3][0].info.flags & ptamd.PT_FLAG_ENV_MIS
        alone = []
        for (sc, ctx), f in zip(made, frames):
            say(f"[gpu] single render {sc.info.width}x{sc.info.height} x {f}")
            alone.append(single(sc, ctx, f))
        counted = [sc.render_params(flags=sc.info.flags | ptamd.PT_FLAG_COUNT_TRAVERSAL) for sc, _ in made]
        alone_counted = [ctx.render(p, 0, f) for (_, ctx), p, f in zip(made, counted, frames)]
        say("[gpu] batch: next week, moving spheres, cornell, meshes")
        st = ptamd.render_batch([c for _, c in made], [s.render_params() for s, _ in made], frames)
        for (sc, ctx), (acc, n, _), f in zip(made, alone, frames):
            got, m = ctx.read_accum(sc.info.width, sc.info.height)
            assert m == n == f
            assert np.array_equal(got, acc)
        assert st.extension_rays == sum(a[2].extension_rays for a in alone)
        assert st.shadow_rays == sum(a[2].shadow_rays for a in alone)
        assert st.exhausted_paths == sum(a[2].exhausted_paths for a in alone)
        say("[gpu] the same batch, counting traversal")
        stc = ptamd.render_batch([c for _, c in made], counted, frames)
        for k in ("extension_rays", "shadow_rays", "exhausted_paths", "node_fetches", "prim_fetches",
                  "instance_fetches", "shadow_node_fetches", "shadow_prim_fetches", "shadow_instance_fetches"):
            assert getattr(stc, k) == sum(getattr(a, k) for a in alone_counted), k
        assert stc.node_fetches > 0 and stc.shadow_node_fetches > 0
        for (sc, ctx), (acc, _, _) in zip(made, alone):
            got, _ = ctx.read_accum(sc.info.width, sc.info.height)
            assert np.array_equal(got, acc)
    finally:
        for _, ctx in made:
            ctx.close()


def test_batch_graph_replay_is_bit_identical(scenes):
    """PT_FLAG_GRAPH in params[0]: the whole batch is captured once and
    replayed; every replay (with other renders of the same contexts in
    between) leaves each accumulation equal to the direct batch's."""
    frames = [5, 3, 4, 7]
    ctxs = [c for _, c in scenes]
    params = [s.render_params() for s, _ in scenes]
    ptamd.render_batch(ctxs, params, frames, stats=False)
    direct = [c.read_accum(s.info.width, s.info.height)[0] for s, c in scenes]
    gparams = list(params)
    gparams[0] = scenes[0][0].render_params(flags=scenes[0][0].info.flags | ptamd.PT_FLAG_GRAPH)
    for rep in range(3):
        say(f"[gpu] batch graph replay {rep}")
        st = ptamd.render_batch(ctxs, gparams, frames)
        for (s, c), d in zip(scenes, direct):
            got, m = c.read_accum(s.info.width, s.info.height)
            assert m == frames[scenes.index((s, c))]
            assert np.array_equal(got, d)
        assert st.paths == sum(s.info.width * s.info.height * f for (s, _), f in zip(scenes, frames))
        # eager work on the contexts between replays
        single(scenes[1][0], scenes[1][1], 2)
        ctxs[0].sync()


def test_batch_autotune_picks_a_shape_with_identical_results(scenes):
    frames = [5, 3, 4, 7]
    ctxs = [c for _, c in scenes]
    params = [s.render_params() for s, _ in scenes]
    ptamd.render_batch(ctxs, params, frames, stats=False)
    direct = [c.read_accum(s.info.width, s.info.height)[0] for s, c in scenes]
    say("[gpu] batch autotune")
    r = ptamd.autotune_batch(ctxs, params, [1, 1, 1, 1])
    # six sub-batch sizes (8 Mi .. 256 Mi, at most the automatic batch size,
    # which a 288 GB device does not reach) x sort x graph
    assert r.candidates in (4 * k for k in range(1, 7)), r.candidates
    assert r.paths_per_batch in [1 << k for k in range(23, 29)]
    assert r.ms_per_frame > 0
    tuned = list(params)
    tuned[0] = scenes[0][0].render_params(flags=scenes[0][0].info.flags | ptamd.PT_FLAG_TUNED)
    ptamd.render_batch(ctxs, tuned, frames, stats=False)
    for (s, c), d in zip(scenes, direct):
        assert np.array_equal(c.read_accum(s.info.width, s.info.height)[0], d)


# the 64-scene workload's largest mesh scenes (1.31 M and 1.18 M triangles),
# its largest sphere scene (4478 spheres) and five more, at full resolution
C3_SUBSET = [49, 63, 44, 33, 20, 1, 11, 7]


def test_workload_subset_matches_oracle():
    """Eight scenes of the 64-scene workload at their own resolutions, 2
    frames each, in one batch: on each scene's stratified 1/64 pixel subset
    the accumulation equals the oracle's per-pixel sums bit for bit, and
    its relMSE against the scene's FP64 reference is reported."""
    import ptref
    from pathlib import Path
    from ptamd.metrics import accum_at, relmse, stratified_pixels

    made = [make(ptamd.PT_SCENE_BATCH, 0, 0, i) for i in C3_SUBSET]
    fx = Path(__file__).resolve().parent / "golden" / "fp64" / "c3_batch.npz"
    refs = np.load(fx) if fx.exists() else None
    try:
        frames = [2] * len(made)
        say("[gpu] 8 scenes of the batch workload, 2 frames each")
        st = ptamd.render_batch([c for _, c in made], [s.render_params() for s, _ in made], frames)
        assert st.paths == sum(s.info.width * s.info.height * 2 for s, _ in made)
        for i, (sc, ctx) in zip(C3_SUBSET, made):
            W, H = sc.info.width, sc.info.height
            acc, m = ctx.read_accum(W, H)
            assert m == 2
            pix = stratified_pixels(W, H)
            say(f"[cpu] oracle, scene {i} ({W}x{H}, {len(pix)} pixels)")
            sums, _ = ptref.OracleScene(sc.desc).render_pixels(sc.camera, sc.render_params(), 0, 2, pix, threads=16)
            assert np.array_equal(accum_at(acc, pix, W, H), sums), i
            if refs is not None:
                rp = refs[f"pix_{i:02d}"]
                e = relmse(accum_at(acc, rp, W, H) / 2, refs[f"mean_{i:02d}"])
                say(f"[gpu] scene {i}: relMSE at 2 spp vs FP64 reference {e:.4f}")
                assert np.isfinite(e)
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
