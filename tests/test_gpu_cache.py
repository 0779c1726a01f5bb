"""GPU checks of the radiance-cache network kernels (BF16 MFMA forward,
backward through the activations, split-K weight-gradient GEMM, Adam)
against the NumPy restatement in oracle/ptref_mlp.py, through the C ABI.

Tolerances: the oracle rounds operands to BF16 at the same points as the
kernels and accumulates in float64; the kernels accumulate in FP32 in MFMA
order, which moves a BF16 rounding of a hidden activation by one BF16 ulp
(2^-8 relative) now and then.  Outputs: max error <= 2e-2 of the output
scale, mean <= 1e-3; gradients: per-layer relative norm error <= 2e-2 and
cosine >= 0.999; Adam: the closed-form float32 update to 1e-6 relative.
"""
import sys

import numpy as np
import pytest

import ptamd
import ptref_mlp as M

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ptamd.Context(0)
    yield c
    c.close()


def features(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1, 1, (n, 96)).astype(np.float32)
    f[:, 83:] = 0.0  # padding of the renderer's 83 features
    return M.to_bf16_bits(f)


def layer_slices():
    out, off = [], 0
    for o, i in M.SIZES:
        out.append(slice(off, off + o * i + o))
        off += o * i + o
    return out


def test_forward_matches_oracle(ctx):
    say("[gpu] cache forward")
    ctx.cache_init(1)
    p = ctx.cache_params()
    f = features(1000, 1)  # not a multiple of 16: tail lanes are masked
    y = ctx.cache_infer(f)
    yr, _ = M.forward(p, M.from_bf16_bits(f), bf16_ops=True)
    err = np.abs(y - yr)
    scale = np.abs(yr).max()
    assert err.max() <= 2e-2 * scale and err.mean() <= 1e-3 * scale, (err.max(), err.mean(), scale)


def test_gradients_match_oracle(ctx):
    say("[gpu] cache gradients")
    ctx.cache_init(2)
    p = ctx.cache_params()
    n = 2500  # padded to 2528: three split-K chunks, a partial last one
    f = features(n, 2)
    t = np.random.default_rng(3).uniform(-1, 2, (n, 3)).astype(np.float32)
    loss, g = ctx.cache_train(f, t, lr=0.0)
    lref, gref = M.loss_and_grad(p, M.from_bf16_bits(f), t, bf16_ops=True)
    assert abs(loss - lref) <= 1e-3 * lref
    for sl in layer_slices():
        a, b = g[sl].astype(np.float64), gref[sl]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        assert rel <= 2e-2 and cos >= 0.999, (sl, rel, cos)
    # lr = 0 leaves the parameters unchanged
    np.testing.assert_array_equal(ctx.cache_params(), p)


def test_adam_matches_closed_form(ctx):
    say("[gpu] cache Adam")
    ctx.cache_init(3)
    p0 = ctx.cache_params()
    f = features(512, 4)
    t = np.random.default_rng(5).uniform(0, 1, (512, 3)).astype(np.float32)
    m = np.zeros(M.NPARAM, np.float32)
    v = np.zeros(M.NPARAM, np.float32)
    p = p0
    for step in (1, 2, 3):
        _, g = ctx.cache_train(f, t, lr=1e-3)
        p, m, v = M.adam(p, m, v, g, step, lr=1e-3)
        got = ctx.cache_params()
        np.testing.assert_allclose(got, p, rtol=1e-6, atol=1e-9)
        p = got


def test_training_fits_a_smooth_function(ctx):
    say("[gpu] cache training run")
    ctx.cache_init(4)
    rng = np.random.default_rng(6)
    n = 8192
    x = rng.uniform(-1, 1, (n, 96)).astype(np.float32)
    x[:, 83:] = 0
    # positive radiance-like targets over two decades
    t = np.stack([np.exp(np.sin(3 * x[:, 0]) + x[:, 1]), 0.5 + x[:, 2] * x[:, 3], 2.0 + x[:, :8].mean(1)],
                 1).astype(np.float32)
    f = M.to_bf16_bits(x)
    first, _ = ctx.cache_train(f, t, lr=1e-3)
    for _ in range(300):
        last, _ = ctx.cache_train(f, t, lr=1e-3)
    assert last < 0.1 * first, (first, last)
    y = ctx.cache_infer(f)
    rel = M.rel_l2_terms(y, t)[0].sum(1).mean()
    assert rel < 0.1 * first, (rel, first)


# ------------------------------------------------ the cache in the renderer
def cornell(ctx, W, H):
    sc = ptamd.Scene(ptamd.PT_SCENE_CORNELL, W, H)
    ctx.set_scene(sc.desc)
    ctx.build_accel()
    ctx.set_camera(sc.camera)
    return sc


def render_image(ctx, params, f0, n):
    st = ctx.render(params, f0, n)
    acc, frames = ctx.read_accum(params.width, params.height)
    return acc / frames, st


def relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01)))


def test_training_records_leave_paths_unchanged(ctx):
    """With no queries (query depth beyond max depth) the cache only records
    training vertices: the image is bit-identical to the plain render, and
    each training path's first vertex (L = 0, T = 1) has the path's own
    result as its target."""
    say("[gpu] cache records without queries")
    W = H = 64
    sc = cornell(ctx, W, H)
    ctx.cache_init(5)
    ctx.cache_configure(query_depth=1000, train_fraction=0.5, batch=4096, steps=2)
    plain = sc.render_params(flags=ptamd.PT_FLAG_KEEP_PATHS, frames_per_batch=1)
    cached = sc.render_params(flags=ptamd.PT_FLAG_KEEP_PATHS | ptamd.PT_FLAG_RADIANCE_CACHE, frames_per_batch=1)
    a, _ = render_image(ctx, plain, 0, 3)
    b, st = render_image(ctx, cached, 0, 3)
    assert np.array_equal(a, b)
    assert st.cache_queries == 0 and st.cache_train_vertices > 0 and st.cache_steps == 2
    rgb, sig, bounces = ctx.read_paths(W, H)
    feats, vals, slots, total = ctx.cache_records(0)
    assert total == len(slots) > 0.4 * W * H
    first = np.all(vals[:, 0:3] == 0, 1) & np.all(vals[:, 3:6] == 1, 1)
    assert first.sum() > 0.4 * W * H
    want = np.maximum(rgb[slots[first]], 0)
    np.testing.assert_allclose(vals[first, 6:9], want, rtol=1e-5, atol=1e-6)
    f = M.from_bf16_bits(feats)
    np.testing.assert_array_equal(f[:, 83:], 0)
    s2c2 = f[:, 0] ** 2 + f[:, 1] ** 2  # first frequency of x: sin^2 + cos^2
    assert np.all(np.abs(s2c2 - 1) < 2e-2)
    albedos = {(0.65, 0.05, 0.05), (0.73, 0.73, 0.73), (0.12, 0.45, 0.15), (0.0, 0.0, 0.0)}
    rd = {tuple(float(v) for v in np.round(x, 2)) for x in f[:, 77:80].astype(np.float64)}
    assert rd <= albedos, rd


def test_queries_end_paths_and_training_converges(ctx):
    say("[gpu] cache queries + online training")
    W = H = 128
    sc = cornell(ctx, W, H)
    ctx.cache_init(6)
    # 10 % training paths: about 5k training vertices per 128^2 frame, four
    # Adam steps of up to 4096 of them per frame
    ctx.cache_configure(query_depth=2, train_fraction=0.1, batch=4096, steps=4)
    params = sc.render_params(flags=ptamd.PT_FLAG_KEEP_PATHS | ptamd.PT_FLAG_RADIANCE_CACHE, frames_per_batch=1)
    losses, judged, fell = [], [], []
    for k in range(60):
        st = ctx.render(params, k, 1)
        losses.append(st.cache_loss)
        judged.append(round(st.cache_judge, 4))
        fell.append(st.cache_fallbacks + st.cache_plain_batches)  # frames not answered by the network
    say(f"[gpu] judge per frame {judged}")
    say(f"[gpu] fallbacks per frame {fell}")
    say(f"[gpu] loss first {np.round(losses[:5], 4)}, last {np.round(losses[-5:], 4)}")
    # the relative L2 loss of single-path targets keeps their variance (its
    # minimiser is their mean), so the loss itself levels off; the judge's
    # error ratio (answers against ending the paths) falls as the network
    # learns, and the judge passes the network's answers once it has
    assert all(np.isfinite(losses))
    assert np.mean(judged[-10:]) < 0.2 * np.mean(judged[:5]), judged
    assert sum(fell[-20:]) <= 2, (fell, judged)
    # one more frame with the judge's answers always taken: the records of
    # a batch whose queries were answered
    ctx.cache_judge_configure(max_ratio=1e30)
    try:
        ctx.render(params, 60, 1)  # its judge passes the network: the next frame asks it
        st = ctx.render(params, 61, 1)
    finally:
        ctx.cache_judge_configure()
    assert st.cache_fallbacks == 0 and st.cache_plain_batches == 0
    _, _, bounces = ctx.read_paths(W, H)
    _, _, tslots, _ = ctx.cache_records(0)
    qf, qv, qslots, nq = ctx.cache_records(1)
    assert nq == st.cache_queries > 0.3 * W * H
    # the batch's answers (inference over the queries where the shading
    # kernel left them) equal a dense inference of the same features
    np.testing.assert_array_equal(qv[:, 6:9], ctx.cache_infer(qf))
    assert np.unique(qslots).size == nq
    # paths that were not training paths end by the third vertex; every query
    # ends such a path there, and the ones that end there without a query
    # are the rays that left the box through its open front
    others = np.setdiff1d(np.arange(W * H), tslots)
    assert np.all(bounces[others] <= 3)
    assert np.all(bounces[qslots] == 3)
    third = others[bounces[others] == 3]
    assert np.setdiff1d(qslots, third).size == 0
    escaped = np.setdiff1d(third, qslots).size
    say(f"[gpu] {nq} queries, {escaped} third-segment escapes of {third.size}")
    assert escaped < 0.3 * third.size, (escaped, third.size)


def test_judge_fallback_renders_the_plain_image(ctx):
    """A new network is held: until the judge passes it, batches render
    plain (no queries; training paths still train).  A judge that passes
    nothing (max ratio 0) therefore keeps every batch plain: the
    accumulation equals the plain render's bit for bit, and no batch needs
    rendering twice.  A network that passed once and then fails is caught
    on the batch that asked it: that batch is rendered again by the plain
    tracer (a fallback) and the next one is held."""
    say("[gpu] cache judge: held batches and fallback to the plain tracer")
    W = H = 96
    sc = cornell(ctx, W, H)
    ctx.cache_init(8)
    ctx.cache_configure(query_depth=2, train_fraction=0.03, batch=1024, steps=2)
    cached = sc.render_params(flags=ptamd.PT_FLAG_RADIANCE_CACHE, frames_per_batch=2)
    plain = sc.render_params(flags=0, frames_per_batch=2)
    a, _ = render_image(ctx, plain, 0, 6)
    try:
        ctx.cache_judge_configure(max_ratio=0.0, rows=2048)
        b, st = render_image(ctx, cached, 0, 6)
        assert st.cache_fallbacks == 0 and st.cache_plain_batches == 3, (st.cache_fallbacks, st.cache_plain_batches)
        assert st.cache_queries == 0 and st.cache_judge >= 0.0
        assert np.array_equal(a, b)
        # pass the network once (any finite ratio), then fail it
        ctx.cache_judge_configure(max_ratio=1e30, rows=2048)
        st = ctx.render(cached, 0, 2)  # held (the judge above failed it); judged: passes
        assert st.cache_plain_batches == 1 and st.cache_fallbacks == 0
        ctx.cache_judge_configure(max_ratio=0.0, rows=2048)
        st = ctx.render(cached, 2, 4)  # batch 1 asks and falls back, batch 2 is held
        assert st.cache_fallbacks == 1 and st.cache_plain_batches == 1, (st.cache_fallbacks, st.cache_plain_batches)
        b, _ = ctx.read_accum(W, H)
    finally:
        ctx.cache_judge_configure()
    assert np.array_equal(a, b / 6)


def test_forced_divergence_falls_back_to_plain(ctx):
    """A learning rate of 1e3 drives the network's outputs out of range
    within a batch or two: the judge refuses the answers of the batch that
    asked the diverged network (rendered again by the plain tracer) and
    holds the later ones, and every such frame is the plain tracer's, path
    for path, with a finite accumulation."""
    say("[gpu] cache judge: forced divergence")
    W = H = 96
    sc = cornell(ctx, W, H)
    ctx.cache_init(9)
    ctx.cache_configure(query_depth=2, train_fraction=0.05, batch=1024, steps=4)
    cached = sc.render_params(flags=ptamd.PT_FLAG_RADIANCE_CACHE | ptamd.PT_FLAG_KEEP_PATHS, frames_per_batch=1)
    plain = sc.render_params(flags=ptamd.PT_FLAG_KEEP_PATHS, frames_per_batch=1)
    # a sane network that the judge has passed: frame 1 asks it
    ctx.cache_judge_configure(max_ratio=1e30)
    ctx.render(cached, 0, 1)
    ctx.cache_judge_configure()
    ctx.cache_configure(query_depth=2, train_fraction=0.05, batch=1024, steps=4, lr=1e3)
    fell, held, judged = [], [], []
    for f in range(1, 9):
        st = ctx.render(cached, f, 1)
        fell.append(st.cache_fallbacks)
        held.append(st.cache_plain_batches)
        judged.append(st.cache_judge)
        if st.cache_fallbacks or st.cache_plain_batches:
            rgb_c, sig_c, _ = ctx.read_paths(W, H)
            acc, _ = ctx.read_accum(W, H)
            assert np.all(np.isfinite(acc))
            ctx.render(plain, f, 1)  # frame f alone (f > 0 continues an accumulation, harmless here)
            rgb_p, sig_p, _ = ctx.read_paths(W, H)
            assert np.array_equal(sig_c, sig_p) and np.array_equal(rgb_c, rgb_p, equal_nan=True)
    say(f"[gpu] fallbacks per frame {fell}, held {held}, judge {judged}")
    # the first frame or the second asks the network that diverged in training
    first = next((i for i, x in enumerate(fell) if x), None)
    assert first is not None and first <= 1, (fell, held, judged)
    assert all(a + b == 1 for a, b in zip(fell[first:], held[first:])), (fell, held, judged)
    assert sum(fell[first + 1:]) == 0, (fell, held, judged)  # held after the fallback, not rendered twice
    assert all(not (0.0 <= j < 1.0) for j in judged[first:]), judged


def test_equal_time_gain_over_plain_tracer(ctx):
    """relMSE against a converged image of the plain tracer (4096 spp): the
    cache form, trained online on the same view, against the plain form,
    both rendering 128 frames in one 128-frame batch (2 Mi paths).

    The plain tracer is unbiased, so its variance falls as 1 / time; its
    error is scaled by t_plain / t_cache to the cache's GPU time (training
    and the host round trip for the record counts included).  The scaling
    also shrinks the reference's own noise, which stays in the plain
    error at any time, so it errs in the plain tracer's favour.  The cache
    must be better per sample and at equal time."""
    say("[gpu] cache vs plain at equal time")
    W = H = 128
    fpb = 128
    sc = cornell(ctx, W, H)
    plain = sc.render_params(flags=0, frames_per_batch=fpb)
    ctx.render(plain, 0, 0)  # a cleared image at this resolution
    ref, _ = render_image(ctx, plain, 1000, 4096)
    img_p, st_p = render_image(ctx, plain, 0, fpb)
    ctx.cache_init(7)
    # 16 Adam steps of 4096 vertices per 128-frame batch (a third of the
    # batch's training vertices); 4 steps leave the network under-trained
    ctx.cache_configure(query_depth=2, train_fraction=0.03, batch=4096, steps=16)
    cached = sc.render_params(flags=ptamd.PT_FLAG_RADIANCE_CACHE, frames_per_batch=fpb)
    # warm-up training on the same view: 32 batches of 8 frames, 16 Adam
    # steps each (the accumulation is cleared by the render below)
    ctx.render(sc.render_params(flags=ptamd.PT_FLAG_RADIANCE_CACHE, frames_per_batch=8), 5000, 256)
    img_c, st_c = render_image(ctx, cached, 0, fpb)
    assert st_c.cache_fallbacks == 0 and st_c.cache_plain_batches == 0, st_c.cache_judge
    e_p, e_c = relmse(img_p, ref), relmse(img_c, ref)
    e_eq = e_p * st_p.gpu_ms / st_c.gpu_ms
    say(f"[gpu] {fpb} spp: plain {st_p.gpu_ms:.2f} ms relMSE {e_p:.5f}; cache {st_c.gpu_ms:.2f} ms "
        f"relMSE {e_c:.5f}; plain at the cache's time {e_eq:.5f}")
    assert e_c < e_p
    assert e_c <= e_eq


def test_mesh_orbit_cache_records_targets_queries_and_gain():
    """The radiance-cache workload at full size: the 2^20-triangle mesh
    scene under the HDRI at 1920x1080 on the orbit camera, 1 spp per frame,
    the cache training online for frames 0-50.  Then, at the frozen view of
    frame 50:
      * training records: targets are (L_final - L) / T of the slot's own
        path result (the unbiased radiance the network learns), features
        are finite and bounded;
      * queries: one per path that reached the query vertex, answered with
        the network's dense inference of the same features;
      * gain: at 64 spp (frames 0-63 from a cleared image, the cache still
        training), the cache's relMSE against the oracle's 16384-spp FP64
        reference of that view is below the plain tracer's: the paths that
        reach the query vertex through glass no longer carry the sun
        caustics (peak 5e4) whose rare samples decide the plain relMSE.  The
        trimmed relMSE (without the 0.1 % worst pixels) is reported."""
    from pathlib import Path

    from ptamd.metrics import accum_at, orbit_camera, relmse_trimmed
    from ptamd.metrics import relmse as rel

    say("[gpu] mesh orbit with the radiance cache")
    fx = Path(__file__).resolve().parent / "golden" / "fp64" / "c4_orbit_f050.npz"
    sc = ptamd.Scene(ptamd.PT_SCENE_MESHES)
    W, H = sc.info.width, sc.info.height
    with ptamd.Context(0) as ctx:
        ctx.set_scene(sc.desc)
        ctx.build_accel()
        ctx.cache_init(0)
        ctx.cache_configure()
        base = sc.info.flags
        cached = sc.render_params(flags=base | ptamd.PT_FLAG_RADIANCE_CACHE | ptamd.PT_FLAG_KEEP_PATHS,
                                  frames_per_batch=1)
        plain = sc.render_params(flags=base, frames_per_batch=1)
        judged, fell = [], []
        for f in range(51):
            ctx.set_camera(orbit_camera(sc.info, f))
            # frames 49 and 50 take any finite answers, so that frame 50 is
            # answered and its query records exist
            ctx.cache_judge_configure(max_ratio=1e30 if f >= 49 else 1.0)
            st = ctx.render(cached, f, 1)
            judged.append(round(st.cache_judge, 3))
            fell.append(st.cache_fallbacks + st.cache_plain_batches)
        ctx.cache_judge_configure()
        say(f"[gpu] judge per frame {judged}")
        say(f"[gpu] fallbacks {sum(fell)} of 51; queries {st.cache_queries}, "
            f"training vertices {st.cache_train_vertices}")
        # most camera paths end on their own before the query vertex (sky,
        # ground lit by sun and sky): about 2 % of them are answered
        assert st.cache_queries > 0.01 * W * H and st.cache_train_vertices > 0.005 * W * H
        # records of the last frame (its queries answered unless it fell back)
        rgb, _, _ = ctx.read_paths(W, H)
        tf, tv, tslots, nt = ctx.cache_records(0)
        assert nt == len(tslots) == st.cache_train_vertices
        f64 = M.from_bf16_bits(tf).astype(np.float64)
        assert np.all(np.isfinite(f64)) and np.all(np.abs(f64) <= 1.0 + 1e-2)
        L0, T0, tgt = tv[:, 0:3], tv[:, 3:6], tv[:, 6:9]
        ok = np.all(T0 > 0, 1) & np.all(np.isfinite(rgb[tslots]), 1)
        want = np.where(T0 > 0, (rgb[tslots] - L0) / np.where(T0 > 0, T0, 1), 0)
        want = np.where((want > 0) & (want < 3e38), want, 0)
        np.testing.assert_allclose(tgt[ok], want[ok].astype(np.float32), rtol=1e-5, atol=1e-6)
        assert fell[-1] == 0, (fell, judged)
        qf, qv, qslots, nq = ctx.cache_records(1)
        assert nq == st.cache_queries and np.unique(qslots).size == nq
        np.testing.assert_array_equal(qv[:, 6:9], ctx.cache_infer(qf))
        # the frozen view of frame 50: 64 spp each from a cleared image
        z = np.load(fx)
        ctx.set_camera(orbit_camera(sc.info, 50))
        ctx.render(plain, 0, 64)
        acc_p, _ = ctx.read_accum(W, H)
        st = ctx.render(cached, 0, 64)
        acc_c, _ = ctx.read_accum(W, H)
    img_p = accum_at(acc_p, z["pix"], W, H) / 64
    img_c = accum_at(acc_c, z["pix"], W, H) / 64
    e_p, e_c = rel(img_p, z["mean"]), rel(img_c, z["mean"])
    t_p, t_c = relmse_trimmed(img_p, z["mean"]), relmse_trimmed(img_c, z["mean"])
    say(f"[gpu] frozen view 50 at 64 spp: relMSE plain {e_p:.4f}, cache {e_c:.4f}; trimmed plain {t_p:.5f}, "
        f"cache {t_c:.5f}; of 64 cache frames {st.cache_fallbacks} fell back, {st.cache_plain_batches} held")
    assert np.isfinite(e_c) and np.isfinite(t_c) and e_c < e_p
