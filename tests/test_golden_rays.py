"""The committed golden ray sets (tests/golden/rays_*.npz, written by
tests/golden/make_ray_fixture.py): 4096 random rays through the
random-spheres scene and the 2^20-triangle mesh scene with the oracle's
closest hits and traversal counts.

CPU: the oracle still produces them (a regression check of the restatement
and of the scene generators), and on the sphere scene its BVH answer equals
brute force over every primitive.  GPU: the closest-hit query kernel
(pt_trace_closest_counted) reproduces instance, primitive, t and the
per-ray node, primitive and instance counts bit for bit."""
from pathlib import Path

import numpy as np
import pytest

import ptamd
import ptref

GOLDEN = Path(__file__).resolve().parent / "golden"
SCENES = {"random_spheres": ptamd.PT_SCENE_RANDOM_SPHERES, "meshes": ptamd.PT_SCENE_MESHES}


def fixture(name):
    z = np.load(GOLDEN / f"rays_{name}.npz")
    return z["rays"].view(ptamd.RAY_DTYPE).reshape(-1), z["inst"], z["prim"], z["t"], z["counts"]


@pytest.mark.parametrize("name", list(SCENES))
def test_oracle_reproduces_golden_rays(name):
    rays, inst, prim, t, counts = fixture(name)
    sc = ptamd.Scene(SCENES[name])  # owns the description's arrays
    orc = ptref.OracleScene(sc.desc)
    hits, cnt = orc.trace(rays, counted=True)
    assert np.array_equal(hits["inst"], inst) and np.array_equal(hits["prim"], prim)
    assert np.array_equal(hits["t"].view(np.uint32), t.view(np.uint32))
    assert np.array_equal(cnt, counts)
    if name == "random_spheres":
        brute = orc.trace(rays, brute=True)
        assert np.array_equal(brute["inst"], inst) and np.array_equal(brute["prim"], prim)
        assert np.array_equal(brute["t"].view(np.uint32), t.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(SCENES))
def test_gpu_matches_golden_rays(name):
    rays, inst, prim, t, counts = fixture(name)
    sc = ptamd.Scene(SCENES[name])
    print(f"[gpu] golden rays {name}", flush=True)
    with ptamd.Context(0) as ctx:
        ctx.set_scene(sc.desc)
        ctx.build_accel()
        hits, cnt = ctx.trace_closest(rays, counted=True)
    assert np.array_equal(hits["inst"], inst) and np.array_equal(hits["prim"], prim)
    assert np.array_equal(hits["t"].view(np.uint32), t.view(np.uint32))
    assert np.array_equal(cnt, counts)
