"""Pins the NumPy restatement of the radiance-cache network: its backward
pass against float64 central finite differences, Adam against the closed
form of its first step, and the BF16 rounding against exact cases."""
import numpy as np

import ptref_mlp as M


def random_problem(n=32, seed=0):
    rng = np.random.default_rng(seed)
    params = np.zeros(M.NPARAM)
    off = 0
    for o, i in M.SIZES:
        params[off:off + o * i] = rng.uniform(-1, 1, o * i) * np.sqrt(6 / i)
        off += o * i
        params[off:off + o] = rng.uniform(-0.1, 0.1, o)
        off += o
    feats = rng.uniform(-1, 1, (n, 96))
    targets = rng.uniform(-1, 2, (n, 3))
    return params, feats, targets


def test_backward_matches_finite_differences():
    """The gradient holds each denominator yh^2 + eps constant, so the
    finite differences are of the loss with the denominators frozen at the
    base point."""
    params, feats, targets = random_problem()
    _, g = M.loss_and_grad(params, feats, targets, bf16_ops=False)
    y0, _ = M.forward(params, feats, bf16_ops=False)
    den = M.rel_l2_terms(y0, targets)[2]
    rng = np.random.default_rng(1)
    # every bias and 300 random weights, spread over all layers
    idx = np.concatenate([rng.choice(M.NPARAM, 300, replace=False),
                          np.array([6144, 6200, 10304, 14464, 18624, 18880, 18882])])
    h = 1e-6
    for p in idx:
        a = params.copy()
        a[p] += h
        b = params.copy()
        b[p] -= h
        num = (M.loss_and_grad(a, feats, targets, False, den)[0] -
               M.loss_and_grad(b, feats, targets, False, den)[0]) / (2 * h)
        assert abs(num - g[p]) <= 1e-6 * max(1.0, abs(g[p])), (p, num, g[p])


def test_loss_definition():
    params, feats, targets = random_problem(n=8, seed=2)
    y, _ = M.forward(params, feats, bf16_ops=False)
    loss, _ = M.loss_and_grad(params, feats, targets, bf16_ops=False)
    r = np.expm1(y)
    assert np.isclose(loss, np.sum((r - targets) ** 2 / (r ** 2 + 0.01)) / 8)


def test_relative_loss_closed_forms():
    """Scalar cases of the relative L2 term and its derivative in y."""
    ln = np.log
    cases = [  # y, t, loss, dloss/dy
        (0.0, 1.0, 1.0 / 0.01, 2 * (0 - 1) * 1 / 0.01),
        (ln(3.0), 2.0, 0.0, 0.0),
        (ln(2.0), 0.0, 1.0 / 1.01, 2 * 1 * 2 / 1.01),
        (ln(0.5), 0.5, 1.0 / 0.26, 2 * (-0.5 - 0.5) * 0.5 / 0.26),
    ]
    for y, t, want_l, want_d in cases:
        l, d, _ = M.rel_l2_terms(np.array([y]), np.array([t]))
        assert np.isclose(l[0], want_l, rtol=1e-12, atol=1e-12), (y, t, l)
        assert np.isclose(d[0], want_d, rtol=1e-12, atol=1e-12), (y, t, d)
    # above the clamp the radiance is expm1(20) and the slope is taken there
    l, d, _ = M.rel_l2_terms(np.array([30.0]), np.array([0.0]))
    big = np.expm1(20.0)
    assert np.isclose(l[0], big * big / (big * big + 0.01)) and np.isclose(d[0], 2 * big * (big + 1) / (big * big + 0.01))
    assert np.isnan(M.rel_l2_terms(np.array([np.nan]), np.array([1.0]))[0][0])
    # answers: clamped, never negative, NaN answers 0
    np.testing.assert_array_equal(M.radiance(np.array([-3.0, np.nan, 0.0])), [0.0, 0.0, 0.0])
    assert M.radiance(np.array([50.0]))[0] == np.expm1(20.0)


def test_relative_loss_minimiser_is_the_mean():
    """With the denominators held constant, the summed gradient over noisy
    targets of one input vanishes exactly where the answer is their mean
    (an unbiased cache), and points towards it elsewhere."""
    rng = np.random.default_rng(7)
    t = rng.exponential(2.0, 1000) * (rng.uniform(size=1000) < 0.1) * 50  # sparse fireflies
    m = t.mean()
    for y, sign in ((np.log1p(m), 0), (np.log1p(0.5 * m), -1), (np.log1p(2 * m), 1)):
        g = M.rel_l2_terms(np.full(t.size, y), t)[1].sum()
        if sign == 0:
            assert abs(g) <= 1e-9 * np.abs(M.rel_l2_terms(np.full(t.size, y), t)[1]).sum()
        else:
            assert np.sign(g) == sign


def test_adam_first_step_is_lr_times_sign():
    """With bias correction, step 1 gives m_hat = g and v_hat = g^2, so the
    update is -lr g / (|g| + eps)."""
    g = np.float32([3.0, -0.5, 1e-3, 0.0, -2e4])
    p0 = np.float32([1, 1, 1, 1, 1])
    p, m, v = M.adam(p0, np.zeros(5, np.float32), np.zeros(5, np.float32), g, 1, lr=1e-3)
    want = p0 - np.float32(1e-3) * g / (np.abs(g) + np.float32(1e-8))
    np.testing.assert_allclose(p, want, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=1e-4)  # 1 - 0.999 in float32


def test_bf16_rounding():
    x = np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -2.5, np.inf, 1e-40, 3.0e38])
    bits = M.to_bf16_bits(x)
    assert bits[0] == 0x3F80
    assert bits[1] == 0x3F80          # tie -> even
    assert bits[2] == 0x3F81          # above the tie -> up
    assert M.from_bf16_bits(bits[3]) == -2.5
    assert np.isinf(M.from_bf16_bits(bits[4]))
    assert np.isnan(M.from_bf16_bits(M.to_bf16_bits(np.float32([np.nan]))))[0]
    assert np.isinf(M.from_bf16_bits(bits[6])) or M.from_bf16_bits(bits[6]) > 3e38
