"""Writes the golden ray sets: 4096 random rays per scene (origins uniform
in the scene's box, directions uniform on the sphere, numpy MT19937 seeded
1) through the random-spheres scene and the 2^20-triangle mesh scene, with
the oracle's closest hits (instance, primitive, t) and per-ray traversal
counts (nodes, primitives, instances entered).

  python tests/golden/make_ray_fixture.py

Output: tests/golden/rays_<scene>.npz (rays, inst, prim, t, counts).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "optix-path-tracer_amd"))
sys.path.insert(0, str(REPO / "oracle"))

import ptamd  # noqa: E402

N = 4096
SCENES = {"random_spheres": ptamd.PT_SCENE_RANDOM_SPHERES, "meshes": ptamd.PT_SCENE_MESHES}


def random_rays(lo, hi, n, seed=1):
    g = np.random.Generator(np.random.MT19937(seed))
    rays = np.zeros(n, ptamd.RAY_DTYPE)
    rays["org"] = (lo + (hi - lo) * g.random((n, 3))).astype(np.float32)
    d = g.normal(size=(n, 3))
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"] = 1e-4
    rays["tmax"] = np.inf
    return rays


def main():
    import ptref

    for name, cfg in SCENES.items():
        sc = ptamd.Scene(cfg)
        orc = ptref.OracleScene(sc.desc)
        lo, hi = fixed_box(name)
        rays = random_rays(np.array(lo), np.array(hi), N)
        hits, cnt = orc.trace(rays, counted=True)
        np.savez_compressed(HERE / f"rays_{name}.npz", rays=rays, inst=hits["inst"], prim=hits["prim"], t=hits["t"],
                            counts=cnt, lo=np.array(lo, np.float32), hi=np.array(hi, np.float32))
        print(f"{name}: {N} rays, {(hits['inst'] >= 0).mean():.2f} hit, mean nodes {cnt[:, 0].mean():.2f}")


def fixed_box(name):
    """Ray-origin box per scene (the region the scene's geometry fills)."""
    return {"random_spheres": ((-12.0, 0.0, -12.0), (12.0, 3.0, 12.0)),
            "meshes": ((-20.0, -12.0, -20.0), (20.0, 12.0, 20.0))}[name]


if __name__ == "__main__":
    main()
