"""Pins the elementary functions shared by the oracle and the kernels
(oracle/ptref_libm.hpp, optix-path-tracer_amd/csrc/pt_libm.hpp) against
numpy's double-precision functions: accuracy in float32 ulps over the
domains the path tracer uses, plus exact special values."""
import numpy as np

import ptref


def ulp_err(got, ref):
    sp = np.spacing(np.abs(ref.astype(np.float32))).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / sp


def check(fn, x, ref, max_ulp, y=None, floor=1e-6):
    got = ptref.libm(fn, x, y)
    keep = np.abs(ref) > floor
    e = ulp_err(got[keep], ref[keep])
    assert e.max() <= max_ulp, (fn, e.max(), x[keep][e.argmax()])
    return got


def test_sin_cos():
    x = np.linspace(-50, 50, 200001).astype(np.float32)
    xd = x.astype(np.float64)
    for fn, f in (("sin", np.sin), ("cos", np.cos)):
        got = check(fn, x, f(xd), 2.0, floor=1e-3)
        assert np.abs(got - f(xd)).max() < 1.2e-7
    # 2*pi*u for u in [0, 1): the argument of every disk / azimuth sample
    u = np.arange(0, 1 << 16, dtype=np.float32) / np.float32(1 << 16)
    phi = np.float32(2.0) * np.float32(np.pi) * u
    for fn, f in (("sin", np.sin), ("cos", np.cos)):
        assert np.abs(ptref.libm(fn, phi) - f(phi.astype(np.float64))).max() < 1.2e-7


def test_tan_atan():
    x = np.linspace(-1.5, 1.5, 100001).astype(np.float32)
    check("tan", x, np.tan(x.astype(np.float64)), 3.0, floor=1e-3)
    x = np.concatenate([np.linspace(-100, 100, 100001), np.geomspace(1e-6, 1e6, 5001),
                        -np.geomspace(1e-6, 1e6, 5001)]).astype(np.float32)
    check("atan", x, np.arctan(x.astype(np.float64)), 2.0)


def test_atan2_quadrants():
    rng = np.random.default_rng(0)
    a = rng.normal(size=100000).astype(np.float32)
    b = rng.normal(size=100000).astype(np.float32)
    check("atan2", a, np.arctan2(a.astype(np.float64), b.astype(np.float64)), 3.0, y=b)
    got = ptref.libm("atan2", np.float32([0, 1, 0, 1, -1]), np.float32([1, 0, -1, 1, -1]))
    np.testing.assert_array_equal(got[:3], np.float32([0, np.pi / 2, np.pi]))
    np.testing.assert_array_max_ulp(got[3:], np.float32([np.pi / 4, -3 * np.pi / 4]), maxulp=1)


def test_asin_acos():
    x = np.linspace(-1, 1, 100001).astype(np.float32)
    xd = x.astype(np.float64)
    check("asin", x, np.arcsin(xd), 3.0)
    check("acos", x, np.arccos(xd), 3.0)
    np.testing.assert_array_equal(ptref.libm("acos", np.float32([1, -1, 0])),
                                  np.float32([0, np.pi, np.pi / 2]))
    np.testing.assert_array_equal(ptref.libm("asin", np.float32([0, 1, -1])),
                                  np.float32([0, np.pi / 2, -np.pi / 2]))


def test_pow():
    """pow is 2^(y log2 x) in float: the rounding of y log2 x, an absolute
    error of a few 2^-24 |z| in z = log2 x^y, becomes a relative error of
    the result, so the bound grows with |z|: within 2.5 (1 + |z|) ulp, over
    the general domain and over the Ashikhmin-Shirley lobe's (x in (0, 1],
    y in [10, 1000]) and samplers' (y in (0, 1])."""
    rng = np.random.default_rng(1)
    a = rng.uniform(1e-4, 10, 100000).astype(np.float32)
    b = rng.uniform(-8, 8, 100000).astype(np.float32)
    x = rng.uniform(0.05, 1, 100000).astype(np.float32)
    for xs, ys in ((a, b), (x, (10 ** rng.uniform(1, 3, 100000)).astype(np.float32)),
                   (x, rng.uniform(1e-3, 1, 100000).astype(np.float32))):
        ref = np.power(xs.astype(np.float64), ys.astype(np.float64))
        ok = (ref > 1e-30) & (ref < 1e30)
        e = ulp_err(ptref.libm("pow", xs, ys)[ok], ref[ok])
        z = np.abs(np.log2(ref[ok]))
        assert np.all(e <= 2.5 * (1 + z)), (e / (1 + z)).max()
    np.testing.assert_array_equal(ptref.libm("pow", np.float32([3, 1, 0, 2, 4]), np.float32([0, 7.5, 2, 10, 0.5])),
                                  np.float32([1, 1, 0, 1024, 2]))
    u = rng.uniform(0, 1, 100000).astype(np.float32)
    ref5 = u.astype(np.float64) ** 5
    ok = ref5 > 1e-30
    assert ulp_err(ptref.libm("pow5", u)[ok], ref5[ok]).max() <= 3.0
    np.testing.assert_array_equal(ptref.libm("pow5", np.float32([0, 1, 0.5, 2])), np.float32([0, 1, 1 / 32, 32]))
