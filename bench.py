"""Benchmark: Mrays/s of the per-sample render path on one MI355X.

Default run (`python bench.py`): one JSON line whose headline is the
triangle-mesh scene (8 displaced icospheres, 2^20 triangles, on a ground
quad under an importance-sampled 2048x1024 HDRI with a sun lobe; 1920x1080,
256 samples per pixel, max depth 12, material sort), the largest single-GPU
configuration; beside it, in `configs`, the Cornell box with light + BRDF
MIS (512x512, 64 spp, depth 16, Russian roulette from depth 3) and the
random-spheres scene (400x267, 16 spp, depth 50).  One step renders every
sample of the image (pt_render, frames 0..spp-1) with the scene,
acceleration structure and path queues resident in HBM before the timed
region, in the launch shape pt_autotune picked for the view beforehand
(paths per batch, material sort, graph replay; reported as `tuned_shape`).

  value          = (closest-hit + shadow queries) / wall time of the timed steps
  relmse_vs_fp64 = relMSE of the step's image against a converged FP64
                   render of the oracle (16384 spp, tests/golden/fp64) on its
                   stratified 1/64 pixel subset
  roofline       = the dominant kernel: algorithmic bytes per launch (the
                   traversal's node, primitive and instance fetches at their
                   stored sizes, counted by a traversal-counting pass over the
                   identical deterministic rays, plus the path state each
                   launch reads and writes) / its average launch time, HIP
                   events on the render stream; `traffic` = HBM bytes per
                   launch from rocprofv3 FETCH_SIZE + WRITE_SIZE, each scaled
                   by the same counter over a plain device copy
  cpu_baseline   = the scalar CPU restatement (oracle) on the host's cores,
                   on a bounded sample of the same workload (named in
                   `sample`), in Mrays/s

Other workloads: `--workload cache` (the learned radiance cache on the mesh
scene's orbit: equal-error time of cache and plain tracer against the FP64
references of the frozen views), `--workload batch` (64 uneven scenes in
one pt_render_batch), `--workload meshes|cornell|random_spheres` (one plain
configuration alone).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "optix-path-tracer_amd"))
FIXTURES = REPO / "tests" / "golden" / "fp64"

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# radiance-cache network 96 -> 64 -> 64 -> 64 -> 64 -> 3: multiply-adds of
# the forward pass (6144 + 3 x 4096 + 192), of the backward pass through the
# activations (layers 2-5: 3 x 4096 + 192) and of the weight gradients
MLP_FWD_FLOP = 2 * 18624
MLP_TRAIN_FLOP = 2 * (18624 + 12480 + 18624)
CPU_THREADS_MAX = 16  # the host-core share of one GPU on the measurement box
# stored record bytes per primitive kind (pt_layout.hpp)
PRIM_BYTES = {0: 16, 1: 32, 2: 32, 3: 48}
# algorithmic bytes of the LBVH build per primitive: bounds (32 B write +
# read), Morton code and index (8 B write, 4 passes of 8 B read + 8 B write
# in the radix sort), hierarchy and refit (64 B), BVH4 collapse (16 B)
BUILD_BYTES_PER_PRIM = 130


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration(mib: int):
    """A plain device-to-device copy of known size (reads and writes `mib`
    MiB once), for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copied {mib} MiB three times")


def host_cpu():
    """Model name and core counts of the host."""
    model = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"model": model, "hardware_concurrency": os.cpu_count(), "available": avail}


def load_profile(name):
    """Per-kernel figures measured with rocprofv3 on this workload
    (profiles/<name>.json): HBM bytes per launch and counter summaries."""
    f = REPO / "profiles" / f"{name}.json"
    try:
        return json.loads(f.read_text())["kernels"]
    except Exception:
        return {}


def load_fixture(name):
    f = FIXTURES / f"{name}.npz"
    if not f.exists():
        return None
    d = np.load(f)
    return d["pix"], d["mean"].astype(np.float64), int(d["spp"])


def relmse_at(ctx, W, H, fixture):
    """(relMSE, trimmed relMSE) of the context's accumulated image against a
    converged reference on its pixel subset (trimmed: without the 0.1 % of
    pixels of largest error)."""
    from ptamd.metrics import accum_at, relmse, relmse_trimmed

    pix, ref, _ = fixture
    acc, frames = ctx.read_accum(W, H)
    img = accum_at(acc, pix, W, H) / max(frames, 1)
    return relmse(img, ref), relmse_trimmed(img, ref)


# (scene config, reference image, record bytes of the scene's primitives,
#  workload name, CPU-baseline sample)
PLAIN = {
    "meshes": ("PT_SCENE_MESHES", "c2_meshes", 48, "meshes_2p20_tris_hdri_envmis_{W}x{H}_{spp}spp_depth{d}",
               "subset"),
    "cornell": ("PT_SCENE_CORNELL", "c1_cornell", 32, "cornell_box_mis_{W}x{H}_{spp}spp_depth{d}_rr{rr}", "full"),
    "random_spheres": ("PT_SCENE_RANDOM_SPHERES", "c0_random_spheres", 16,
                       "random_spheres_{W}x{H}_{spp}spp_depth{d}_sky", "full"),
}


def cpu_baseline(scene, params, W, H, spp, how, fixture):
    """The oracle on the host cores over a bounded sample of the workload:
    `full` renders every frame of the step, `subset` the full sample count on
    the stratified 1/64 pixel subset of the FP64 reference."""
    sys.path.insert(0, str(REPO / "oracle"))
    import ptref

    cpu = host_cpu()
    threads = max(1, min(CPU_THREADS_MAX, cpu["available"]))
    orc = ptref.OracleScene(scene.desc)
    t1 = time.perf_counter()
    if how == "full":
        _, _, st = orc.render(scene.camera, params, 0, spp, threads=threads)
        rays = int(st[1]) + int(st[2])
        sample = f"the whole step: {W}x{H}, frames 0-{spp - 1} ({W * H * spp} paths, {rays} rays)"
    else:
        pix = fixture[0]
        _, st = orc.render_pixels(scene.camera, params, 0, spp, pix, threads=threads)
        rays = int(st[1]) + int(st[2])
        sample = (f"the stratified 1/64 pixel subset ({len(pix)} pixels) at the full {spp} spp "
                  f"({len(pix) * spp} paths, {rays} rays)")
    dt = time.perf_counter() - t1
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{sample}, {dt:.1f} s", "host": cpu}


def plain_config(name, args, with_cpu=True):
    """Measure one plain-tracer configuration; returns its result dict."""
    import ptamd

    cfg, fixname, prim_b, wname, cpu_how = PLAIN[name]
    scene = ptamd.Scene(getattr(ptamd, cfg))
    info = scene.info
    W, H = info.width, info.height
    spp = args.spp or info.spp
    params = scene.render_params()
    ctx = ptamd.Context(0)
    ctx.set_scene(scene.desc)
    acc = ctx.build_accel()
    ctx.set_camera(scene.camera)
    log(f"[bench] {name}: {acc.num_primitives} primitives, {acc.num_nodes} nodes, {acc.num_instances} instances, "
        f"build {acc.build_ms:.2f} ms")

    # launch shape: pt_autotune's pick for this view over the whole step
    # (paths per batch, material sort, graph replay; untimed)
    t_tune = time.perf_counter()
    if args.shape:
        # a fixed shape of direct launches (profiler runs: rocprofv3's
        # kernel trace has crashed inside graph capture)
        fpb, so = (int(x) for x in args.shape.split(","))
        tune = ptamd.TuneResult(frames_per_batch=fpb, material_sort=so, graph=0, candidates=0,
                                paths_per_batch=fpb * W * H)
    else:
        tune = ctx.autotune(params, spp)
    t_tune = time.perf_counter() - t_tune
    shape = {"frames_per_batch": tune.frames_per_batch, "paths_per_batch": tune.paths_per_batch,
             "material_sort": bool(tune.material_sort), "graph": bool(tune.graph), "candidates": tune.candidates,
             "tuning_s": round(t_tune, 1)}
    log(f"[bench] {name}: {'fixed' if args.shape else 'tuned'} shape {shape}")
    # direct launches of the same shape, for per-kernel HIP-event timings
    sort_bit = ptamd.PT_FLAG_MATERIAL_SORT if tune.material_sort else 0
    direct = scene.render_params(flags=(info.flags & ~ptamd.PT_FLAG_MATERIAL_SORT) | sort_bit,
                                 frames_per_batch=tune.frames_per_batch)
    tuned = direct if args.shape else scene.render_params(flags=info.flags | ptamd.PT_FLAG_TUNED)
    # traversal-counting pass (identical rays: the render is deterministic)
    cst = ctx.render(scene.render_params(flags=(info.flags & ~ptamd.PT_FLAG_MATERIAL_SORT) | sort_bit |
                                         ptamd.PT_FLAG_COUNT_TRAVERSAL, frames_per_batch=tune.frames_per_batch), 0, spp)
    log(f"[bench] {name}: counted pass {cst.extension_rays} closest-hit, {cst.shadow_rays} shadow rays")
    for _ in range(args.warmup):
        ctx.render(tuned, 0, spp)
    ctx.sync()
    steps = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        steps.append(ctx.render(tuned, 0, spp))
        log(f"[bench] {name} step {k}: {steps[-1].gpu_ms:.2f} ms")
    ctx.sync()
    wall = time.perf_counter() - t0
    assert all(s.extension_rays == cst.extension_rays and s.shadow_rays == cst.shadow_rays for s in steps)
    kst = [ctx.render(direct, 0, spp) for _ in range(2)][-1]  # kernel timings (the image is the same)
    fixture = load_fixture(fixname)
    err, err_t = relmse_at(ctx, W, H, fixture) if fixture else (None, None)
    # the build again, timed alone (the scene and its structures are resident)
    builds = [ctx.build_accel().build_ms for _ in range(3)]
    ctx.close()

    rays = sum(s.extension_rays + s.shadow_rays for s in steps)
    ext, paths, exh = cst.extension_rays, cst.paths, cst.exhausted_paths
    # k_shade: per shaded path 64 B state (32 B at depth 0) + 16 B hit read,
    # 64 B survivor or 20 B result written; the shadow ray's traversal
    shade_bytes = (ext * (64 + 16) - paths * 32 + (ext - paths + exh) * 64 + (paths - exh) * 20 +
                   cst.shadow_node_fetches * 128 + cst.shadow_prim_fetches * prim_b +
                   cst.shadow_instance_fetches * 64)
    # k_extend: 32 B ray + 16 B hit per ray, 16 B result header per path at
    # depth 0, the traversal's fetches
    ext_bytes = (ext * (32 + 16) + paths * 16 + cst.node_fetches * 128 + cst.prim_fetches * prim_b +
                 cst.instance_fetches * 64)
    shade_ms = kst.shade_ms
    ext_ms = kst.extend_ms
    prof = load_profile(f"traffic_{name}")

    def roof(kernel, key, bytes_step, ms, launches, extra):
        achieved = bytes_step / (ms * 1e-3) / 1e9
        p = prof.get(key, {})
        r = {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": p.get("hbm_bytes_per_launch"),
             "algorithmic_bytes_per_launch": round(bytes_step / launches), "avg_launch_ms": round(ms / launches, 4)}
        if p.get("hbm_bytes_per_launch"):
            r["traffic_GBs"] = round(p["hbm_bytes_per_launch"] / (ms / launches * 1e-3) / 1e9, 1)
        r["node_bytes"] = 128  # BVH4 node: 6 x 4 child slab planes + 4 child refs + 4 counts, fp32
        r["prim_bytes"] = prim_b
        for k in ("valu_busy", "occupancy_waves_per_simd", "l2_hit", "wait_frac"):
            if k in p:
                r[k] = p[k]
        r.update(extra)
        return r

    r_shade = roof("k_shade (shading + shadow-ray BVH4 traversal + compaction)", "k_shade", shade_bytes, shade_ms,
                   kst.shade_launches, {"per_shadow_ray": {
                       "nodes": round(cst.shadow_node_fetches / max(1, cst.shadow_rays), 3),
                       "prims": round(cst.shadow_prim_fetches / max(1, cst.shadow_rays), 3),
                       "instances": round(cst.shadow_instance_fetches / max(1, cst.shadow_rays), 3)}})
    r_ext = roof("k_extend (closest-hit BVH4 traversal; camera paths started at depth 0)", "k_extend", ext_bytes,
                 ext_ms, kst.extend_launches, {"per_ray": {
                     "nodes": round(cst.node_fetches / ext, 3), "prims": round(cst.prim_fetches / ext, 3),
                     "instances": round(cst.instance_fetches / ext, 3)}})
    dominant, other = (r_ext, r_shade) if ext_ms >= shade_ms else (r_shade, r_ext)
    for r in (dominant, other):
        # measured traffic far below the algorithmic bytes: the structure
        # stays in L2 / MALL and the kernel is bound by instruction issue;
        # the VALU fraction of the SIMDs' cycles (rocprofv3 SQ_INSTS_VALU x 4
        # cycles per wave64 instruction over 1024 SIMDs x the dispatch's
        # cycles) is the figure that binds it
        if r.get("valu_busy") and r.get("traffic") and r["traffic"] < 0.25 * r["algorithmic_bytes_per_launch"]:
            r["issue"] = {"bound": "valu", "achieved": r["valu_busy"], "peak": 1.0,
                          "unit": "fraction of SIMD cycles issuing VALU", "frac": r["valu_busy"],
                          "note": "HBM traffic is {:.0%} of the algorithmic bytes (cache-resident structure)".format(
                              r["traffic"] / r["algorithmic_bytes_per_launch"])}
    build_ms = min(builds)
    out = {
        "metric": "Mrays/s",
        "value": round(rays / wall / 1e6, 2),
        "unit": "Mrays/s",
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "config": {"workload": wname.format(W=W, H=H, spp=spp, d=info.max_depth, rr=info.rr_min_depth),
                   "width": W, "height": H, "spp": spp, "max_depth": info.max_depth,
                   "rr_min_depth": info.rr_min_depth,
                   "material_sort": bool(info.flags & ptamd.PT_FLAG_MATERIAL_SORT),
                   "env_mis": bool(info.flags & ptamd.PT_FLAG_ENV_MIS)},
        "samples_per_s": round(args.steps * W * H * spp / wall),
        "relmse_vs_fp64": None if err is None else round(err, 6),
        "relmse_trimmed_vs_fp64": None if err_t is None else round(err_t, 6),
        "relmse_reference": None if fixture is None else
        f"oracle FP64 build, {fixture[2]} spp, {len(fixture[0])} pixels (stratified 1/64)",
        "rays_per_step": {"closest_hit": cst.extension_rays, "shadow": cst.shadow_rays},
        "nan_paths_per_step": cst.nan_paths,
        "kernel_ms_per_step": {"extend": round(ext_ms, 3), "shade": round(shade_ms, 3),
                               "render_events": round(sum(s.gpu_ms for s in steps) / len(steps), 3),
                               "direct_launches_events": round(kst.gpu_ms, 3)},
        "tuned_shape": shape,
        "build": {"ms": round(build_ms, 3), "primitives": acc.num_primitives, "nodes": acc.num_nodes,
                  "algorithmic_GBs": round(BUILD_BYTES_PER_PRIM * acc.num_primitives / (build_ms * 1e-3) / 1e9, 1)},
        "roofline": dominant,
        "roofline_other": other,
    }
    if with_cpu and not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(scene, params, W, H, spp, cpu_how, fixture)
        except Exception as e:  # the baseline is reported, never the target
            log(f"[bench] cpu baseline failed: {e}")
            out["cpu_baseline"] = None
    return out


def default_workload(args):
    head = plain_config("meshes", args)
    extra = {}
    for name in ("cornell", "random_spheres"):
        extra[name] = plain_config(name, args)
    out = {"metric": "Mrays/s", "value": head["value"], "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic"}
    out.update({k: v for k, v in head.items() if k not in out and k != "metric"})
    # the 64-scene batch (its own line: bench.py --workload batch)
    try:
        extra["batch"] = batch_workload(args)
    except Exception as e:
        log(f"[bench] batch workload failed: {e}")
        extra["batch"] = None
    out["configs"] = extra
    print(json.dumps(out), flush=True)


def single_workload(args):
    r = plain_config(args.workload, args)
    out = {"metric": "Mrays/s", "value": r["value"], "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic"}
    out.update({k: v for k, v in r.items() if k not in out})
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------ cache

def _time_to(curve, e, extrapolate):
    """Time at which a relMSE-vs-time curve [(t, relmse)] first reaches
    error e, interpolated linearly in log t / log relmse between
    checkpoints; beyond the last checkpoint an unbiased estimator's
    relMSE falls as 1/t (`extrapolate`), a biased one's is not assumed to
    fall at all (None)."""
    prev = None
    for t, r in curve:
        if r <= e:
            if prev is None:  # reached at the first checkpoint
                return t * r / e if extrapolate else t
            (t0, r0) = prev
            if r0 == r:
                return t
            s = (math.log(e) - math.log(r0)) / (math.log(r) - math.log(r0))
            return math.exp(math.log(t0) + s * (math.log(t) - math.log(t0)))
        prev = (t, r)
    if extrapolate:
        t, r = curve[-1]
        return t * r / e
    return None


def cache_workload(args):
    """The mesh scene at 1920x1080 on an orbiting camera (radius 20, one
    turn per 300 frames), 1 spp per frame, the radiance cache training
    online (3 % of paths traced to the end as training paths, queries after
    2 bounces).  At the frozen views of frames 50, 100, 200 and 300 the
    plain tracer and the cache tracer (continuing to train) each render up
    to `--eval-spp` frames progressively from a cleared image; relMSE
    against the oracle's FP64 16384-spp reference of that view is taken
    after 1, 2, 4, ... frames, against the wall time of each form (host
    round trips included).  Reported per view: both curves and the time
    the cache needs to reach the plain tracer's final error, divided by the
    plain tracer's time (the target is at most 1/3)."""
    import ptamd
    from ptamd.metrics import orbit_camera

    frozen = [f for f in (50, 100, 200, 300) if f <= args.frames]
    scene = ptamd.Scene(ptamd.PT_SCENE_MESHES)
    W, H = scene.info.width, scene.info.height
    ctx = ptamd.Context(0)
    ctx.set_scene(scene.desc)
    ctx.build_accel()
    ctx.cache_init(0)
    ctx.cache_configure()
    base = scene.info.flags
    plain = scene.render_params(flags=base, frames_per_batch=1)
    cached = scene.render_params(flags=base | ptamd.PT_FLAG_RADIANCE_CACHE, frames_per_batch=1)

    def timed(p, f0, n):
        ctx.sync()
        t = time.perf_counter()
        st = ctx.render(p, f0, n)
        ctx.sync()
        return (time.perf_counter() - t) * 1e3, st

    def curve(p, fixture):
        """[(wall ms, relMSE, trimmed relMSE, frames)] after 1, 2, 4, ...
        frames from a cleared image"""
        pts, t, done = [], 0.0, 0
        n = 1
        while done < args.eval_spp:
            k = min(n, args.eval_spp) - done
            ms, _ = timed(p, done, k)
            t += ms
            done += k
            e, et = relmse_at(ctx, W, H, fixture)
            pts.append((round(t, 3), round(e, 6), round(et, 6), done))
            n *= 2
        return pts

    # warm-up: kernels loaded, pools sized (frame 0 of both forms)
    ctx.set_camera(orbit_camera(scene.info, 0))
    timed(plain, 0, 1)
    views, anim = [], {"plain_ms": 0.0, "cache_ms": 0.0, "fallbacks": 0, "held": 0, "frames": 0}
    mlp = {"train_ms": 0.0, "infer_ms": 0.0, "train_samples": 0, "infer_samples": 0}
    for f in range(args.frames + 1):
        ctx.set_camera(orbit_camera(scene.info, f))
        ms, st = timed(cached, f, 1)
        anim["cache_ms"] += ms
        anim["fallbacks"] += st.cache_fallbacks
        anim["held"] += st.cache_plain_batches
        anim["frames"] += 1
        mlp["train_ms"] += st.cache_train_ms
        mlp["infer_ms"] += st.cache_infer_ms
        mlp["train_samples"] += st.cache_train_samples
        mlp["infer_samples"] += st.cache_infer_samples
        if f % 25 == 0:
            ms_p, _ = timed(plain, f, 1)
            log(f"[bench] frame {f}: cache {ms:.2f} ms (judge {st.cache_judge:.3f}, {st.cache_queries} queries, "
                f"{st.cache_train_vertices} training vertices), plain {ms_p:.2f} ms")
        if f in frozen:
            fx = load_fixture(f"c4_orbit_f{f:03d}")
            if fx is None:
                raise SystemExit(f"missing reference tests/golden/fp64/c4_orbit_f{f:03d}.npz")
            cp = curve(plain, fx)
            cc = curve(cached, fx)
            v = {"frame": f, "plain": cp, "cache": cc}
            for name, col in (("trimmed", 2), ("full", 1)):
                # the cache's time to reach the plain tracer's error after
                # all its frames, over the plain tracer's time; and the
                # plain tracer's time (1/t beyond its last checkpoint) to
                # reach the cache's final error, over the cache's time
                t_p = cp[-1][0]
                t_c = _time_to([(q[0], q[col]) for q in cc], cp[-1][col], extrapolate=False)
                t_p2 = _time_to([(q[0], q[col]) for q in cp], cc[-1][col], extrapolate=True)
                v[f"time_ratio_at_plain_error_{name}"] = None if t_c is None else round(t_c / t_p, 4)
                v[f"time_ratio_at_cache_error_{name}"] = round(cc[-1][0] / t_p2, 4) if t_p2 else None
            views.append(v)
            log(f"[bench] view {f}: plain {cp[-1]}, cache {cc[-1]}, ratio (trimmed) "
                f"{v['time_ratio_at_plain_error_trimmed']}")
            ctx.set_camera(orbit_camera(scene.info, f))
    ctx.close()
    # MFMA roofline of the network kernels over the animation (HIP events
    # around the training steps and the inference launches)
    flop = mlp["train_samples"] * MLP_TRAIN_FLOP + mlp["infer_samples"] * MLP_FWD_FLOP
    net_ms = mlp["train_ms"] + mlp["infer_ms"]
    achieved = flop / (net_ms * 1e-3) / 1e12 if net_ms > 0 else 0.0
    roofline = {"kernel": "k_mlp_* (BF16 MFMA forward, backward, weight gradients, Adam; inference)",
                "bound": "mfma", "achieved": round(achieved, 3), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / BF16_PEAK_TFLOPS, 6), "traffic": None,
                "train": {"samples": mlp["train_samples"], "ms": round(mlp["train_ms"], 3),
                          "flop_per_sample": MLP_TRAIN_FLOP},
                "infer": {"samples": mlp["infer_samples"], "ms": round(mlp["infer_ms"], 3),
                          "flop_per_sample": MLP_FWD_FLOP}}
    ratios = [v["time_ratio_at_plain_error_trimmed"] for v in views]
    finite = [r for r in ratios if r is not None]
    value = math.exp(sum(math.log(r) for r in finite) / len(finite)) if finite and len(finite) == len(ratios) else None
    out = {
        "metric": "equal-error time, radiance cache / plain (trimmed relMSE vs FP64 oracle)",
        "value": None if value is None else round(value, 4),
        "unit": "ratio",
        "n_gpus": 1,
        "higher_is_better": False,
        "data": "synthetic",
        "dtype": "f32 tracing, bf16 network (fp32 accumulate)",
        "config": {"workload": f"meshes_hdri_orbit_{W}x{H}_1spp_per_frame_cache", "frames": args.frames,
                   "frozen_views": frozen, "eval_spp": args.eval_spp},
        "views": views,
        "animation_ms_per_frame": round(anim["cache_ms"] / max(1, anim["frames"]), 3),
        "animation_fallbacks": anim["fallbacks"],
        "animation_held_plain": anim["held"],
        "roofline": roofline,
        "curves": "[wall ms, relMSE, trimmed relMSE, frames] after 1, 2, 4, ... frames from a cleared image; "
                  "value = geometric mean over the views of time_ratio_at_plain_error_trimmed (trimmed: without "
                  "the 0.1 % of pixels of largest error, whose rare sun caustics decide the full relMSE)",
    }
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------ batch

def batch_workload(args):
    """The 64-scene batch: 32 sphere scenes (50..5000 spheres, sky) and 32
    mesh scenes (14k..1.3M triangles under the HDRI), square resolutions
    256..1024, 4, 16 or 64 spp, every scene's frames rendered by one
    pt_render_batch per step in the launch shape pt_autotune_batch picked
    (paths per sub-batch, material sort, graph replay; tuned on a quarter of
    the frames).  Per-scene relMSE against the oracle's FP64 8192-spp
    references on each scene's stratified 1/256 pixel subset.  Beside it,
    the same scenes rendered one pt_render each, back to back."""
    import ptamd
    from ptamd.metrics import accum_at, relmse, relmse_trimmed

    scenes = [ptamd.Scene(ptamd.PT_SCENE_BATCH, 0, 0, i) for i in range(64)]
    ctxs = [ptamd.Context(0) for _ in scenes]
    t0 = time.perf_counter()
    prims = 0
    build_ms = 0.0
    for sc, ctx in zip(scenes, ctxs):
        ctx.set_scene(sc.desc)
        a = ctx.build_accel()
        ctx.set_camera(sc.camera)
        prims += a.num_primitives
        build_ms += a.build_ms
    log(f"[bench] 64 scenes built ({prims} primitives, {build_ms:.1f} ms of GPU build) in "
        f"{time.perf_counter() - t0:.1f} s")
    params = [sc.render_params() for sc in scenes]
    frames = [sc.info.spp for sc in scenes]
    paths = sum(sc.info.width * sc.info.height * sc.info.spp for sc in scenes)
    t0 = time.perf_counter()
    f0 = scenes[0].info.flags
    tuned = list(params)
    if args.shape:
        # direct launches at the automatic sub-batch size with the given
        # material sort (profiler runs: no graph capture)
        so = int(args.shape.split(",")[-1])
        f0 = (f0 & ~ptamd.PT_FLAG_MATERIAL_SORT) | (ptamd.PT_FLAG_MATERIAL_SORT if so else 0)
        tuned[0] = scenes[0].render_params(flags=f0)
        shape = {"fixed": True, "material_sort": bool(so), "graph": False}
    else:
        tune = ptamd.autotune_batch(ctxs, params, [max(1, f // 4) for f in frames])
        log(f"[bench] batch autotune: {tune.paths_per_batch >> 20} Mi paths per sub-batch, sort "
            f"{tune.material_sort}, graph {tune.graph} ({tune.candidates} candidates, "
            f"{time.perf_counter() - t0:.1f} s)")
        f0 |= ptamd.PT_FLAG_TUNED
        tuned[0] = scenes[0].render_params(flags=f0)
        shape = {"paths_per_batch": tune.paths_per_batch, "material_sort": bool(tune.material_sort),
                 "graph": bool(tune.graph), "candidates": tune.candidates}
    # counted pass (node / primitive fetches per ray) and a direct pass of
    # the same shape for per-kernel HIP-event times; the images are the same
    counted = list(tuned)
    counted[0] = scenes[0].render_params(flags=f0 | ptamd.PT_FLAG_COUNT_TRAVERSAL | ptamd.PT_FLAG_DIRECT)
    cst = ptamd.render_batch(ctxs, counted, frames)
    direct = list(tuned)
    direct[0] = scenes[0].render_params(flags=f0 | ptamd.PT_FLAG_DIRECT)
    kst = [ptamd.render_batch(ctxs, direct, frames) for _ in range(2)][-1]
    st = cst
    rays_step = st.extension_rays + st.shadow_rays
    log(f"[bench] batch: {paths} paths, {st.extension_rays} closest-hit + {st.shadow_rays} shadow rays, "
        f"{kst.gpu_ms:.1f} ms direct ({kst.extend_ms:.1f} closest-hit, {kst.shade_ms:.1f} shading)")
    for _ in range(args.warmup):
        ptamd.render_batch(ctxs, tuned, frames, stats=False)
    ctxs[0].sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ptamd.render_batch(ctxs, tuned, frames, stats=False)
        ctxs[0].sync()
        log(f"[bench] step {k} done")
    wall = time.perf_counter() - t0
    errs, errs_t = [], []
    fx = FIXTURES / "c3_batch.npz"
    refs = np.load(fx) if fx.exists() else None
    if refs is not None:
        for i, (sc, ctx) in enumerate(zip(scenes, ctxs)):
            W, H = sc.info.width, sc.info.height
            acc, n = ctx.read_accum(W, H)
            img, ref = accum_at(acc, refs[f"pix_{i:02d}"], W, H) / n, refs[f"mean_{i:02d}"]
            errs.append(round(relmse(img, ref), 6))
            errs_t.append(round(relmse_trimmed(img, ref), 6))
    # the same work as 64 separate renders (one scene at a time)
    seq_ms = 0.0
    for sc, ctx, p, f in zip(scenes, ctxs, params, frames):
        seq_ms += ctx.render(p, 0, f).gpu_ms
    for ctx in ctxs:
        ctx.close()
    ms = wall * 1e3 / args.steps
    # rooflines of the batch's kernels from the counted pass and the direct
    # pass's kernel times: algorithmic bytes as for the single scenes, with
    # 16 B per primitive fetch (a sphere record; a triangle's is 48 B, so the
    # achieved figures are lower bounds)
    ext, npaths, exh = cst.extension_rays, cst.paths, cst.exhausted_paths
    prim_b = 16
    ext_bytes = (ext * (32 + 16) + npaths * 16 + cst.node_fetches * 128 + cst.prim_fetches * prim_b +
                 cst.instance_fetches * 64)
    shade_bytes = (ext * (64 + 16) - npaths * 32 + (ext - npaths + exh) * 64 + (npaths - exh) * 20 +
                   cst.shadow_node_fetches * 128 + cst.shadow_prim_fetches * prim_b +
                   cst.shadow_instance_fetches * 64)
    prof = load_profile("traffic_batch")

    def roof(kernel, key, b, kms, launches):
        ach = b / (kms * 1e-3) / 1e9
        p = prof.get(key, {})
        r = {"kernel": kernel, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": p.get("hbm_bytes_per_launch"),
             "algorithmic_bytes_per_launch": round(b / max(1, launches)),
             "avg_launch_ms": round(kms / max(1, launches), 4), "node_bytes": 128, "prim_bytes": prim_b}
        for k in ("valu_busy", "occupancy_waves_per_simd", "l2_hit", "wait_frac"):
            if k in p:
                r[k] = p[k]
        return r

    r_ext = roof("k_extend (closest hit, every scene's paths in one launch per bounce; camera paths started at "
                 "depth 0)", "k_extend", ext_bytes, kst.extend_ms, kst.extend_launches)
    r_ext["per_ray"] = {"nodes": round(cst.node_fetches / ext, 3), "prims": round(cst.prim_fetches / ext, 3),
                        "instances": round(cst.instance_fetches / ext, 3)}
    r_shade = roof("k_shade (shading + shadow rays + compaction, every scene's paths)", "k_shade", shade_bytes,
                   kst.shade_ms, kst.shade_launches)
    dominant, other = (r_ext, r_shade) if kst.extend_ms >= kst.shade_ms else (r_shade, r_ext)
    cpu = None
    if not args.no_cpu:
        try:
            cpu = batch_cpu_baseline(scenes)
        except Exception as e:  # the baseline is reported, never the target
            log(f"[bench] batch cpu baseline failed: {e}")
    out = {
        "metric": "Mrays/s",
        "value": round(rays_step * args.steps / wall / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch_64_scenes_32_spheres_50-5000_32_meshes_14k-1.3M_256-1024px_4-64spp",
                   "paths": paths, "primitives": prims},
        "samples_per_s": round(paths * args.steps / wall),
        "rays_per_step": {"closest_hit": st.extension_rays, "shadow": st.shadow_rays},
        "tuned_shape": shape,
        "kernel_ms_per_step": {"extend": round(kst.extend_ms, 3), "shade": round(kst.shade_ms, 3),
                               "direct_launches_events": round(kst.gpu_ms, 3)},
        "roofline": dominant,
        "roofline_other": other,
        "cpu_baseline": cpu,
        "build_ms_total": round(build_ms, 2),
        "one_render_per_scene_ms": round(seq_ms, 2),
        "batch_speedup": round(seq_ms / ms, 3),
        "relmse_vs_fp64": None if not errs else {
            "per_scene": errs, "mean": round(float(np.mean(errs)), 6), "median": round(float(np.median(errs)), 6),
            "max": round(float(np.max(errs)), 6),
            "trimmed_per_scene": errs_t, "trimmed_median": round(float(np.median(errs_t)), 6),
            "reference": "oracle FP64 build, 8192 spp, stratified 1/256 pixel subset per scene; trimmed: without "
                         "the 0.1 % of pixels of largest error"},
    }
    return out


def batch_cpu_baseline(scenes, block=3):
    """The oracle on the host cores over every scene of the batch, each on
    its stratified 1/block^2 pixel subset at its full sample count."""
    sys.path.insert(0, str(REPO / "oracle"))
    import ptref
    from ptamd.metrics import stratified_pixels

    cpu = host_cpu()
    threads = max(1, min(CPU_THREADS_MAX, cpu["available"]))
    rays, npix, dt = 0, 0, 0.0
    for sc in scenes:
        pix = stratified_pixels(sc.info.width, sc.info.height, block)
        orc = ptref.OracleScene(sc.desc)
        t1 = time.perf_counter()
        _, st = orc.render_pixels(sc.camera, sc.render_params(), 0, sc.info.spp, pix, threads=threads)
        dt += time.perf_counter() - t1
        rays += int(st[1]) + int(st[2])
        npix += len(pix)
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"all {len(scenes)} scenes of the batch, stratified 1/{block * block} pixel subsets ({npix} "
                      f"pixels) at each scene's full sample count ({rays} rays), {dt:.1f} s of tracing (the CPU "
                      f"structure builds not timed)", "host": cpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=0, help="override samples per pixel (0: the workload's own)")
    ap.add_argument("--workload", choices=["default", "meshes", "cornell", "random_spheres", "cache", "batch"],
                    default="default",
                    help="default: the mesh scene (headline) plus the Cornell box and random spheres; meshes, "
                         "cornell, random_spheres: one of them alone; cache: the mesh scene's orbit with the "
                         "learned radiance cache against the plain tracer at equal error; batch: 64 sphere and "
                         "mesh scenes of uneven size in one pt_render_batch")
    ap.add_argument("--frames", type=int, default=300, help="cache workload: orbit frames (frozen views at 50, "
                                                            "100, 200, 300)")
    ap.add_argument("--eval-spp", type=int, default=64, help="cache workload: frames rendered per frozen view")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--shape", default="", metavar="FPB,SORT",
                    help="frames per batch and material sort (0/1) of direct launches instead of the autotuner's "
                         "pick (batch workload: only the sort; the automatic sub-batch size)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(samples would shard by frame with one reduction of the FP64 accumulation)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "cache":
        cache_workload(args)
    elif args.workload == "batch":
        print(json.dumps(batch_workload(args)), flush=True)
    elif args.workload == "default":
        default_workload(args)
    else:
        single_workload(args)


if __name__ == "__main__":
    main()
