"""Writes the converged FP64 reference images that relMSE is measured
against: the oracle's double build (Real = double, the 32-bit RNG unchanged)
renders a stratified 1/64 pixel subset of each benchmark view at 16384
samples per pixel, from frame 2^20 on, so that no test or benchmark render
(frames below 2^20) shares a sample with it.

  python tests/golden/make_fp64_refs.py [--spp N] [--threads T] [names...]

Output: tests/golden/fp64/<name>.npz with pix (launch indices x + W*y),
mean (float32, per pixel and channel), and the view's settings.
Views:
  c0_random_spheres   random spheres, 400x267, depth 50
  c1_cornell          Cornell box + MIS, 512x512, depth 16, RR from 3
  c2_meshes           8 displaced icospheres under the HDRI, 1920x1080
  c4_orbit_fNNN       the mesh scene seen from the orbit camera of frame
                      NNN (50, 100, 200, 300): the frozen evaluation views
                      of the radiance-cache workload
  c3_batch            the 64 scenes of the batch workload (32 random-sphere
                      scenes, 32 mesh scenes under the HDRI), each on a
                      stratified 1/256 subset (one pixel per 16 x 16 block)
                      at 8192 spp; one file, arrays pix_NN / mean_NN per
                      scene NN
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "optix-path-tracer_amd"))
sys.path.insert(0, str(REPO / "oracle"))

import ptamd  # noqa: E402
from ptamd.metrics import orbit_camera, stratified_pixels  # noqa: E402

FRAME_BASE = 1 << 20
OUT = HERE / "fp64"

VIEWS = {
    "c0_random_spheres": (ptamd.PT_SCENE_RANDOM_SPHERES, None),
    "c1_cornell": (ptamd.PT_SCENE_CORNELL, None),
    "c2_meshes": (ptamd.PT_SCENE_MESHES, None),
    "c4_orbit_f050": (ptamd.PT_SCENE_MESHES, 50),
    "c4_orbit_f100": (ptamd.PT_SCENE_MESHES, 100),
    "c4_orbit_f200": (ptamd.PT_SCENE_MESHES, 200),
    "c4_orbit_f300": (ptamd.PT_SCENE_MESHES, 300),
}


def view_setup(name):
    """(scene, camera, params) of a reference view."""
    cfg, orbit = VIEWS[name]
    sc = ptamd.Scene(cfg)
    cam = sc.camera if orbit is None else orbit_camera(sc.info, orbit)
    return sc, cam, sc.render_params()


def reference_sums(orc, cam, prm, pix, spp, threads, name=""):
    """Per-pixel sums of the oracle over frames FRAME_BASE + [0, spp), in
    chunks of 1024 frames (each chunk summed in frame order, chunks added in
    order: the result does not depend on the thread count)."""
    sums = np.zeros((len(pix), 3), np.float64)
    stats = np.zeros(4, np.uint64)
    t0 = time.time()
    chunk = 1024
    for f0 in range(0, spp, chunk):
        orc.render_pixels(cam, prm, FRAME_BASE + f0, min(chunk, spp - f0), pix, threads=threads, sums=sums,
                          stats=stats)
        if name:
            print(f"[fp64] {name}: {min(spp, f0 + chunk)}/{spp} spp, {time.time() - t0:.0f} s", flush=True)
    return sums, stats


def make(name, spp, threads):
    import ptref

    sc, cam, prm = view_setup(name)
    W, H = sc.info.width, sc.info.height
    pix = stratified_pixels(W, H)
    orc = ptref.OracleScene(sc.desc, fp64=True)
    t0 = time.time()
    sums, stats = reference_sums(orc, cam, prm, pix, spp, threads, name)
    dt = time.time() - t0
    OUT.mkdir(exist_ok=True)
    cam_vals = np.array(list(cam.origin) + list(cam.lower_left) + list(cam.horizontal) + list(cam.vertical),
                        np.float32)
    np.savez_compressed(OUT / f"{name}.npz", pix=pix, mean=(sums / spp).astype(np.float32),
                        width=W, height=H, spp=spp, frame_base=FRAME_BASE, flags=prm.flags,
                        max_depth=prm.max_depth, rr_min_depth=prm.rr_min_depth, camera=cam_vals,
                        paths=int(stats[0]), rays=int(stats[1] + stats[2]), seconds=dt, threads=threads)
    print(f"[fp64] {name}: {len(pix)} pixels x {spp} spp in {dt:.0f} s", flush=True)


def make_batch(spp, threads, scenes=64):
    import ptref

    out = {}
    t0 = time.time()
    for i in range(scenes):
        sc = ptamd.Scene(ptamd.PT_SCENE_BATCH, 0, 0, i)
        W, H = sc.info.width, sc.info.height
        pix = stratified_pixels(W, H, block=16)
        orc = ptref.OracleScene(sc.desc, fp64=True)
        sums, _ = reference_sums(orc, sc.camera, sc.render_params(), pix, spp, threads)
        out[f"pix_{i:02d}"] = pix
        out[f"mean_{i:02d}"] = (sums / spp).astype(np.float32)
        out[f"size_{i:02d}"] = np.array([W, H], np.uint32)
        print(f"[fp64] c3_batch scene {i}: {W}x{H}, {len(pix)} pixels x {spp} spp, {time.time() - t0:.0f} s",
              flush=True)
    OUT.mkdir(exist_ok=True)
    np.savez_compressed(OUT / "c3_batch.npz", spp=spp, frame_base=FRAME_BASE, block=16, scenes=scenes,
                        seconds=time.time() - t0, threads=threads, **out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=16384)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    ap.add_argument("names", nargs="*", default=list(VIEWS) + ["c3_batch"])
    a = ap.parse_args()
    for n in a.names:
        if n == "c3_batch":
            make_batch(min(a.spp, 8192), a.threads)
        else:
            make(n, a.spp, a.threads)


if __name__ == "__main__":
    main()
