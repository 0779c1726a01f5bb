"""pt_trace_closest_device_noalloc, the entry point that allocates nothing
(its traversal-stack spill lives in the library's per-device scratch
array), against pt_trace_closest_device on the same device-resident rays:
hits bit-identical, on two calls in a row, the second with fewer rays (the
scratch keeps the first call's contents), and no stack overflow.  Sizes
beyond PT_NOALLOC_MAX_RAYS are refused."""
import ctypes as C
import sys

import numpy as np
import pytest

import ptamd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Hip:
    """Device buffers for the test's inputs and outputs (the caller's side)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.bufs = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(max(1, nbytes))) == 0
        self.bufs.append(p)
        return p

    def h2d(self, dst, arr):
        assert self.lib.hipMemcpy(dst, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0

    def d2h(self, arr, src):
        assert self.lib.hipMemcpy(arr.ctypes.data_as(C.c_void_p), src, C.c_size_t(arr.nbytes), 2) == 0

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def random_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, ptamd.RAY_DTYPE)
    rays["org"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"] = 1e-4
    rays["tmax"] = np.inf
    return rays


@pytest.mark.parametrize("cfg,lo,hi", [(ptamd.PT_SCENE_CORNELL, 0.0, 555.0),
                                       (ptamd.PT_SCENE_RANDOM_SPHERES, -12.0, 12.0)])
def test_noalloc_trace_equals_device_trace(cfg, lo, hi):
    sc = ptamd.Scene(cfg, 64, 64)
    hip = Hip()
    try:
        with ptamd.Context(0) as ctx:
            ctx.set_scene(sc.desc)
            ctx.build_accel()
            lib = ctx.lib
            for call, n in enumerate((300_000, 70_001)):
                say(f"[gpu] noalloc trace, call {call}: {n} rays")
                rays = random_rays(n, lo, hi, seed=call)
                d_rays = hip.malloc(rays.nbytes)
                d_a = hip.malloc(n * ptamd.HIT_DTYPE.itemsize)
                d_b = hip.malloc(n * ptamd.HIT_DTYPE.itemsize)
                d_ovf = hip.malloc(4)
                hip.h2d(d_rays, rays)
                assert lib.pt_trace_closest_device(ctx.h, d_rays, n, d_a) == 0
                assert lib.pt_trace_closest_device_noalloc(ctx.h, d_rays, n, d_b, d_ovf) == 0
                ctx.sync()
                a = np.zeros(n, ptamd.HIT_DTYPE)
                b = np.zeros(n, ptamd.HIT_DTYPE)
                ovf = np.zeros(1, np.uint32)
                hip.d2h(a, d_a)
                hip.d2h(b, d_b)
                hip.d2h(ovf, d_ovf)
                assert ovf[0] == 0
                assert (a["inst"] >= 0).mean() > 0.3
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
                # the host-array entry point agrees too
                c = ctx.trace_closest(rays)
                assert np.array_equal(c.view(np.uint8), b.view(np.uint8))
            # beyond the maximum: refused before any launch
            assert lib.pt_trace_closest_device_noalloc(ctx.h, d_rays, (1 << 31) + 1, d_b, None) != 0
            assert "PT_NOALLOC_MAX_RAYS" in lib.pt_last_error(ctx.h).decode()
    finally:
        hip.sync()
        hip.free()
