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
    params, feats, targets = random_problem()
    _, g = M.loss_and_grad(params, feats, targets, bf16_ops=False)
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
        num = (M.loss_and_grad(a, feats, targets, False)[0] - M.loss_and_grad(b, feats, targets, False)[0]) / (2 * h)
        assert abs(num - g[p]) <= 1e-6 * max(1.0, abs(g[p])), (p, num, g[p])


def test_loss_definition():
    params, feats, targets = random_problem(n=8, seed=2)
    y, _ = M.forward(params, feats, bf16_ops=False)
    loss, _ = M.loss_and_grad(params, feats, targets, bf16_ops=False)
    assert np.isclose(loss, np.sum((y - targets) ** 2) / 8)


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
