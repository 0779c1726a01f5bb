"""Benchmark: Mrays/s of the per-sample render path on one MI355X.

Workload (one step): the Cornell box with light + BRDF MIS, 512x512, 64
samples per pixel, max depth 16, Russian roulette from depth 3, rendered
through the C ABI (pt_render, frames 0..63) with the scene, acceleration
structure and path queues resident in HBM before the timed region.

  value    = (closest-hit + shadow queries) / wall time of the timed steps
  roofline = the dominant kernel, k_shade (hit/miss programs, light sample,
             shadow-ray traversal, continuation, compaction of survivors):
             algorithmic bytes = per shaded path 64 B state (32 B at depth
             0, whose throughput and radiance are constants) + 16 B hit
             read, 64 B survivor state or 20 B result written; per shadow ray
             128 B per BVH4 node + record bytes per primitive + 64 B per
             instance entered (counted by a traversal-counting pass over the
             identical deterministic rays).  Material, light and texture
             tables are not counted.  Divided by the kernel's summed duration
             (HIP events on the render stream).
  roofline_extend = the same for k_extend (32 B ray + 16 B hit + traversal;
             the depth-0 launch starts the camera paths: it writes the ray
             instead of reading it, and 16 B of result per path).
  cpu_baseline = the scalar CPU restatement (oracle) on the host cores, on
             all 64 frames on 16 threads, in Mrays/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "optix-path-tracer_amd"))

HBM_PEAK_GBS = 8000.0
# stored record bytes per primitive kind (pt_layout.hpp)
PRIM_BYTES = {0: 16, 1: 32, 2: 32, 3: 48}


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


def load_traffic():
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on
    this workload, scaled by the same counters over a plain copy."""
    f = REPO / "profiles" / "traffic_cornell_512x512_64spp.json"
    try:
        return {k: v["hbm_bytes_per_launch"] for k, v in json.loads(f.read_text())["kernels"].items()}
    except Exception:
        return {}


def orbit_camera(scene, frame, frames_per_turn=300, r=20.0, height=4.0, vfov=45.0):
    """Camera of the radiance-cache workload: orbit of radius r around the
    origin, one turn per `frames_per_turn` frames (frame 0 is the scene's own
    pose at (20, 4, 0))."""
    import math

    import ptamd

    a = 2 * math.pi * frame / frames_per_turn
    o = (r * math.cos(a), height, r * math.sin(a))
    w = [o[k] for k in range(3)]
    n = math.sqrt(sum(x * x for x in w))
    w = [x / n for x in w]
    up = (0.0, 1.0, 0.0)
    u = [up[1] * w[2] - up[2] * w[1], up[2] * w[0] - up[0] * w[2], up[0] * w[1] - up[1] * w[0]]
    n = math.sqrt(sum(x * x for x in u))
    u = [x / n for x in u]
    v = [w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]]
    aspect = scene.info.width / scene.info.height
    hh = math.tan(math.radians(vfov) / 2)
    hw = aspect * hh
    f = 10.0  # focus distance, as the scene camera
    c = ptamd.Camera()
    c.origin[:] = o
    c.lower_left[:] = [o[k] - hw * f * u[k] - hh * f * v[k] - f * w[k] for k in range(3)]
    c.horizontal[:] = [2 * hw * f * u[k] for k in range(3)]
    c.vertical[:] = [2 * hh * f * v[k] for k in range(3)]
    c.u[:] = u
    c.v[:] = v
    c.w[:] = w
    c.lens_radius, c.time0, c.time1 = 0.0, 0.0, 1.0
    return c


def cache_workload(args):
    """Mesh scene, 1 spp per frame on an orbiting camera with the radiance
    cache training online; then, on the frozen view of the last frame, the
    plain tracer and the cache tracer each render 16 frames in one batch and
    their relMSE against a high-spp plain render (the unbiased estimator the
    oracle pins) is compared at the cache's GPU time."""
    import numpy as np

    import ptamd

    W, H = 960, 540
    scene = ptamd.Scene(ptamd.PT_SCENE_MESHES, W, H)
    ctx = ptamd.Context(0)
    ctx.set_scene(scene.desc)
    ctx.build_accel()
    ctx.cache_init(0)
    ctx.cache_configure()
    base = scene.info.flags
    plain = scene.render_params(flags=base)
    cached = scene.render_params(flags=base | ptamd.PT_FLAG_RADIANCE_CACHE, frames_per_batch=1)
    t_plain = t_cache = 0.0
    for f in range(args.frames):
        ctx.set_camera(orbit_camera(scene, f))
        t_plain += ctx.render(plain, f, 1).gpu_ms
        st = ctx.render(cached, f, 1)
        t_cache += st.gpu_ms
        if f % 10 == 0:
            log(f"[bench] frame {f}: cache loss {st.cache_loss:.4f}, {st.cache_queries} queries, "
                f"{st.cache_train_vertices} training vertices")
    ctx.set_camera(orbit_camera(scene, args.frames))
    ref_spp = 1024
    ref_st = ctx.render(plain, 100000, ref_spp)
    ref, _ = ctx.read_accum(W, H)
    ref /= ref_spp
    log(f"[bench] reference {ref_spp} spp in {ref_st.gpu_ms:.1f} ms")
    # the frozen view: both forms render the same number of frames in one
    # batch; the plain tracer is unbiased, so its error is scaled by its
    # time ratio to the cache's GPU time (the reference's own noise is not
    # scaled, which favours the plain tracer)
    spp = 16
    st_p = ctx.render(scene.render_params(flags=base, frames_per_batch=spp), 0, spp)
    img_p, _ = ctx.read_accum(W, H)
    img_p /= spp
    st_c = ctx.render(scene.render_params(flags=base | ptamd.PT_FLAG_RADIANCE_CACHE, frames_per_batch=spp), 0, spp)
    img_c, _ = ctx.read_accum(W, H)
    img_c /= spp
    ctx.close()

    def relmse(a):
        return float(np.mean((a - ref) ** 2 / (ref ** 2 + 0.01)))

    e_p, e_c = relmse(img_p), relmse(img_c)
    e_eq = e_p * st_p.gpu_ms / st_c.gpu_ms
    out = {
        "metric": "equal-time relMSE, radiance cache / plain",
        "value": round(e_c / e_eq, 4),
        "unit": "ratio",
        "n_gpus": 1,
        "higher_is_better": False,
        "data": "synthetic",
        "dtype": "f32 tracing, bf16 network",
        "config": {"workload": f"meshes_hdri_orbit_{W}x{H}_1spp_per_frame_cache", "frames": args.frames},
        "plain": {"spp": spp, "gpu_ms": round(st_p.gpu_ms, 2), "relmse": e_p, "relmse_at_cache_time": e_eq},
        "cache": {"spp": spp, "gpu_ms": round(st_c.gpu_ms, 2), "relmse": e_c},
        "animation_ms_per_frame": {"plain": round(t_plain / args.frames, 3), "cache": round(t_cache / args.frames, 3)},
        "reference_spp": ref_spp,
    }
    print(json.dumps(out), flush=True)


def batch_workload(args):
    """The 64-scene batch: 32 sphere scenes (50..5000 spheres) and 32 mesh
    scenes (10k..2M triangles, sky light), square resolutions 256..1024, 4,
    16 or 64 spp, every scene's frames rendered by one pt_render_batch per
    step.  Beside it, the same scenes rendered one pt_render each, back to
    back (the GPU time the batch saves on small and uneven scenes)."""
    import ptamd

    scenes = [ptamd.Scene(ptamd.PT_SCENE_BATCH, 0, 0, i) for i in range(64)]
    ctxs = [ptamd.Context(0) for _ in scenes]
    t0 = time.perf_counter()
    prims = 0
    for sc, ctx in zip(scenes, ctxs):
        ctx.set_scene(sc.desc)
        a = ctx.build_accel()
        ctx.set_camera(sc.camera)
        prims += a.num_primitives
    log(f"[bench] 64 scenes built ({prims} primitives) in {time.perf_counter() - t0:.1f} s")
    params = [sc.render_params() for sc in scenes]
    frames = [sc.info.spp for sc in scenes]
    paths = sum(sc.info.width * sc.info.height * sc.info.spp for sc in scenes)
    st = ptamd.render_batch(ctxs, params, frames)  # rays of one step (and warm-up)
    rays_step = st.extension_rays + st.shadow_rays
    log(f"[bench] batch: {paths} paths, {st.extension_rays} closest-hit + {st.shadow_rays} shadow rays, "
        f"{st.gpu_ms:.1f} ms")
    for _ in range(args.warmup):
        ptamd.render_batch(ctxs, params, frames, stats=False)
    ctxs[0].sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ptamd.render_batch(ctxs, params, frames, stats=False)
        ctxs[0].sync()
        log(f"[bench] step {k} done")
    wall = time.perf_counter() - t0
    # the same work as 64 separate renders (one scene at a time)
    seq_ms = 0.0
    for sc, ctx, p, f in zip(scenes, ctxs, params, frames):
        seq_ms += ctx.render(p, 0, f).gpu_ms
    for ctx in ctxs:
        ctx.close()
    ms = wall * 1e3 / args.steps
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
        "config": {"workload": "batch_64_scenes_32_spheres_50-5000_32_meshes_10k-2M_256-1024px_4-64spp",
                   "paths": paths, "primitives": prims},
        "samples_per_s": round(paths * args.steps / wall),
        "rays_per_step": {"closest_hit": st.extension_rays, "shadow": st.shadow_rays},
        "one_render_per_scene_ms": round(seq_ms, 2),
        "batch_speedup": round(seq_ms / ms, 3),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=0, help="override samples per pixel (0: the workload's own)")
    ap.add_argument("--workload", choices=["cornell", "meshes", "cache", "batch"], default="cornell",
                    help="cornell: Cornell box 512x512 64 spp (the reported metric); meshes: 8 displaced "
                         "octaspheres (2^20 triangles) under an importance-sampled HDRI, 1920x1080; cache: the "
                         "mesh scene on an orbiting camera with the learned radiance cache, equal-time relMSE; "
                         "batch: 64 sphere and mesh scenes of uneven size in one pt_render_batch")
    ap.add_argument("--frames", type=int, default=60, help="cache workload: animation frames before the evaluation")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-frames", type=int, default=64)
    ap.add_argument("--cpu-threads", type=int, default=16)
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

    import ptamd

    if args.workload == "cache":
        cache_workload(args)
        return
    if args.workload == "batch":
        batch_workload(args)
        return
    meshes = args.workload == "meshes"
    scene = ptamd.Scene(ptamd.PT_SCENE_MESHES if meshes else ptamd.PT_SCENE_CORNELL)
    info = scene.info
    W, H = info.width, info.height
    spp = args.spp or info.spp
    params = scene.render_params()
    ctx = ptamd.Context(0)
    ctx.set_scene(scene.desc)
    acc = ctx.build_accel()
    ctx.set_camera(scene.camera)
    log(f"[bench] accel: {acc.num_nodes} nodes, {acc.num_instances} instances, build {acc.build_ms:.2f} ms")

    # traversal-counting pass (identical rays: the render is deterministic)
    cparams = scene.render_params(flags=info.flags | ptamd.PT_FLAG_COUNT_TRAVERSAL)
    cst = ctx.render(cparams, 0, spp)
    log(f"[bench] counted pass: {cst.extension_rays} closest-hit, {cst.shadow_rays} shadow rays")

    for _ in range(args.warmup):
        ctx.render(params, 0, spp)
    ctx.sync()

    steps = []
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        steps.append(ctx.render(params, 0, spp))
        log(f"[bench] step {k}: {steps[-1].gpu_ms:.2f} ms")
    ctx.sync()
    wall = time.perf_counter() - t0

    rays = sum(s.extension_rays + s.shadow_rays for s in steps)
    ext_rays = sum(s.extension_rays for s in steps)
    mrays = rays / wall / 1e6
    ms_per_step = wall * 1e3 / args.steps
    samples_per_s = args.steps * W * H * spp / wall
    # the counted pass saw exactly the rays of one step
    assert all(s.extension_rays == cst.extension_rays and s.shadow_rays == cst.shadow_rays for s in steps)

    # Roofline of the dominant kernel (k_shade: hit and miss programs, light
    # sampling, the shadow ray's traversal, continuation and compaction) and
    # of the closest-hit traversal kernel (k_extend).
    # record bytes: rects and boxes (32 B) in the Cornell box, triangles
    # (48 B; the one ground rect counted as a triangle) in the mesh scene
    prim_b = PRIM_BYTES[3] if meshes else PRIM_BYTES[1]
    ext, paths, exh = cst.extension_rays, cst.paths, cst.exhausted_paths
    shade_bytes_step = (ext * (64 + 16) - paths * 32 + (ext - paths + exh) * 64 + (paths - exh) * 20 +
                        cst.shadow_node_fetches * 128 + cst.shadow_prim_fetches * prim_b +
                        cst.shadow_instance_fetches * 64)
    ext_bytes_step = (ext * (32 + 16) + paths * 16 + cst.node_fetches * 128 + cst.prim_fetches * prim_b +
                      cst.instance_fetches * 64)
    shade_ms = sum(s.shade_ms for s in steps) / len(steps)
    ext_ms = sum(s.extend_ms for s in steps) / len(steps)
    traffic = {} if meshes else load_traffic()

    def roof(kernel, key, bytes_step, ms, launches, extra):
        achieved = bytes_step / (ms * 1e-3) / 1e9
        t = traffic.get(key)
        return dict({
            "kernel": kernel,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": t,
            "algorithmic_bytes_per_launch": round(bytes_step / launches),
            "avg_launch_ms": round(ms / launches, 4),
        }, **extra)

    roofline = roof("k_shade (shading + shadow-ray BVH4 traversal + compaction)", "k_shade", shade_bytes_step,
                    shade_ms, steps[0].shade_launches, {
                        "shadow_per_ray": {
                            "nodes": round(cst.shadow_node_fetches / max(1, cst.shadow_rays), 3),
                            "prims": round(cst.shadow_prim_fetches / max(1, cst.shadow_rays), 3),
                            "instances": round(cst.shadow_instance_fetches / max(1, cst.shadow_rays), 3)}})
    roofline_extend = roof("k_extend (closest-hit BVH4 traversal; camera paths started at depth 0)", "k_extend",
                           ext_bytes_step, ext_ms,
                           steps[0].extend_launches, {
                               "per_ray": {
                                   "nodes": round(cst.node_fetches / ext, 3),
                                   "prims": round(cst.prim_fetches / ext, 3),
                                   "instances": round(cst.instance_fetches / ext, 3)}})

    cpu = None
    if not args.no_cpu and not meshes:
        try:
            sys.path.insert(0, str(REPO / "oracle"))
            import ptref

            orc = ptref.OracleScene(scene.desc)
            nthreads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
            fr = args.cpu_frames
            t1 = time.perf_counter()
            _, _, st = orc.render(scene.camera, params, 0, fr, threads=nthreads)
            dt = time.perf_counter() - t1
            cpu = {
                "value": round((int(st[1]) + int(st[2])) / dt / 1e6, 3),
                "unit": "Mrays/s",
                "cores": nthreads,
                "kind": "port",
                "sample": f"Cornell 512x512, frames 0-{fr - 1} of 64 ({W * H * fr} paths, "
                          f"{int(st[1]) + int(st[2])} rays, {dt:.1f} s)",
            }
        except Exception as e:  # the baseline is reported, never the target
            log(f"[bench] cpu baseline failed: {e}")

    ctx.close()
    out = {
        "metric": "Mrays/s",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"meshes_2p20_tris_hdri_envmis_{W}x{H}_{spp}spp_depth{info.max_depth}" if meshes else
                         f"cornell_box_mis_{W}x{H}_{spp}spp_depth{info.max_depth}_rr{info.rr_min_depth}"),
            "width": W, "height": H, "spp": spp, "max_depth": info.max_depth,
            "rr_min_depth": info.rr_min_depth, "material_sort": True,
            "env_mis": bool(info.flags & ptamd.PT_FLAG_ENV_MIS),
        },
        "samples_per_s": round(samples_per_s),
        "rays_per_step": {"closest_hit": cst.extension_rays, "shadow": cst.shadow_rays},
        "kernel_ms_per_step": {
            "extend": round(ext_ms, 3),
            "shade": round(shade_ms, 3),
            "render_events": round(sum(s.gpu_ms for s in steps) / len(steps), 3),
        },
        "roofline": roofline,
        "roofline_extend": roofline_extend,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
